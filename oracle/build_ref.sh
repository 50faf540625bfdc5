#!/usr/bin/env bash
# Compile the reference's own CPU core (cpp/src/solver.cpp, cpp/src/interface.cpp, its vendored
# pocketfft header and its pybind11 module cpp/python/bindings.cpp) straight from where the sources
# lie, into oracle/_ref/. Used only by tests (to cross-check the oracle) and by bench.py's CPU
# baseline leg. Nothing from the reference is copied into this repository.
set -euo pipefail
REF=${BWPM_REFERENCE:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
if [ ! -f "$REF/cpp/src/solver.cpp" ]; then
  echo "build_ref: reference sources not found under $REF; skipping" >&2
  exit 0
fi
mkdir -p "$OUT"
SUFFIX=$(python3 -c "import sysconfig;print(sysconfig.get_config_var('EXT_SUFFIX'))")
PYINC=$(python3 -c "import sysconfig;print(sysconfig.get_paths()['include'])")
PBINC=$(python3 -c "import pybind11;print(pybind11.get_include())")
g++ -O3 -std=c++20 -fopenmp -fPIC -shared \
  -I"$REF/cpp/include" -I"$REF/cpp/third_party/pocketfft" -I"$PYINC" -I"$PBINC" \
  "$REF/cpp/python/bindings.cpp" "$REF/cpp/src/solver.cpp" "$REF/cpp/src/interface.cpp" \
  -o "$OUT/_core$SUFFIX.tmp"
mv "$OUT/_core$SUFFIX.tmp" "$OUT/_core$SUFFIX"
echo "build_ref: $OUT/_core$SUFFIX"
