"""CPU checks of the drop-in boundary: the C ABI library loads, exports every entry point
include/bwpm_hip.h declares, maps statuses to the reference's exception types, and the result
codec round-trips the reference's fixture files. No compute call needs a GPU here."""

from __future__ import annotations

import ctypes
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

HEADER = ROOT / "include" / "bwpm_hip.h"


def declared_symbols() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(bwpm_[a-z0-9_]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol() -> None:
    from bwpm_amd import _core

    lib = ctypes.CDLL(_core.LIBRARY)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_reference_module_surface() -> None:
    """bwpm._core names (cpp/python/bindings.cpp:122-180) are present in the binding."""
    import inspect

    from bwpm_amd import _core

    for name in ("transmission_matrix", "reflection_matrix", "longitudinal_wavenumber", "Solver", "CudaSolver",
                 "HAS_CUDA", "cuda_device_count"):
        assert hasattr(_core, name), name
    assert _core.CudaSolver is _core.GpuSolver
    # the positional order of the reference's constructors (bindings.cpp:149-151, 170-172)
    head = ["ny", "nx", "dy", "dx", "dz", "k0", "indices", "labels", "index_before", "mask"]
    assert list(inspect.signature(_core.Solver).parameters)[:12] == head + ["threads", "fluctuation"]
    assert list(inspect.signature(_core.CudaSolver).parameters)[:12] == head + ["device", "fluctuation"]


def test_fft_plan_report_needs_no_device() -> None:
    """The plan tuner's report is host state: empty before any solve, readable without a GPU."""
    from bwpm_amd import _core

    plans = _core.fft_plans()
    assert isinstance(plans, list)
    for p in plans:
        assert p["plan"] in (0, 1) and p["direction"] in ("col", "row")


def test_invalid_arguments_raise_value_error_before_any_device_work() -> None:
    from bwpm_amd import _core

    with pytest.raises(ValueError, match="same shape"):
        _core.transmission_matrix(np.zeros(3), np.zeros(4), 1.0, 1.5, 1.0)
    with pytest.raises(ValueError, match="shape"):
        _core.GpuSolver(4, 4, 1e-4, 1e-4, 1e-4, 1e4, [1.0], np.zeros((2, 4)), 1.0)
    with pytest.raises(ValueError, match="shape"):
        _core.GpuSolver(4, 4, 1e-4, 1e-4, 1e-4, 1e4, [1.0], np.zeros((2, 4, 4)), 1.0,
                        fluctuation=np.zeros((2, 4, 5)))
    with pytest.raises(ValueError, match="precision"):
        _core.GpuSolver(4, 4, 1e-4, 1e-4, 1e-4, 1e4, [1.0], np.zeros((2, 4, 4)), 1.0, precision="fp16")


def test_without_a_gpu_compute_calls_raise_runtime_error() -> None:
    from bwpm_amd import _core

    if _core.gpu_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError):
        _core.transmission_matrix(np.zeros(2), np.zeros(2), 1.0, 1.5, 1.0)
    with pytest.raises(RuntimeError):
        _core.GpuSolver(4, 4, 1e-4, 1e-4, 1e-4, 1e4, [1.0], np.zeros((2, 4, 4)), 1.0)


def test_status_codes_match_the_header() -> None:
    from bwpm_amd import _core

    text = HEADER.read_text()
    for name, value in (("BWPM_OK", _core.OK), ("BWPM_EINVAL", _core.EINVAL), ("BWPM_EDEVICE", _core.EDEVICE),
                        ("BWPM_ENOCONV", _core.ENOCONV), ("BWPM_ENOMEM", _core.ENOMEM)):
        assert re.search(rf"#define {name} {value}\b", text), name
    for name, value in _core.PRECISIONS.items():
        assert f" {value} " in text or f" {value}\n" in text


@pytest.mark.parametrize("path", sorted(GOLDEN.glob("*.result.json")), ids=lambda p: p.name.split(".")[0])
def test_codec_round_trips_the_fixtures(path) -> None:
    from bwpm_amd.results import FIELD_REFERENCE, decode, encode, load_json, to_json

    type_name, data = load_json(path.read_text("utf-8"))
    assert type_name == FIELD_REFERENCE
    again = load_json(to_json(type_name, data))[1]
    assert decode(encode(data))["case"] == data["case"]
    for key, arr in data["expected"].items():
        b = again["expected"][key]
        if isinstance(arr, np.ndarray):
            assert arr.dtype == b.dtype and np.array_equal(arr, b), key
        else:
            assert arr == b


def test_codec_encodes_non_finite_values() -> None:
    from bwpm_amd.results import decode, encode

    a = np.array([1.0, np.nan, np.inf, -np.inf])
    e = encode({"a": a, "c": complex(np.nan, 1.0), "f": float("inf")})
    assert e["a"]["values"][1:] == ["NaN", "Infinity", "-Infinity"]
    d = decode(e)
    assert np.isnan(d["a"][1]) and d["a"][2] == np.inf and d["f"] == np.inf


def test_product_package_never_imports_the_oracle() -> None:
    for path in (ROOT / "bwpm_amd").rglob("*.py"):
        text = path.read_text()
        assert "oracle" not in re.findall(r"^\s*(?:from|import)\s+([a-z_\.]+)", text, re.M), path
        assert "bwpm_oracle" not in text, path
