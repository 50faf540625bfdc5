"""Facade of the GPU engine with the reference's `bwpm.fast` API (src/bwpm/fast.py:61-210).

- `transmission_matrix`, `reflection_matrix`: interface matrices (FP64, device).
- `run_2d(propagator, field, iterations, keep)`: 2D FPWPM (iterations = 0) or BWPM, returns
  `BidirectionalResult` with the per-iteration energy residual.
- `run_3d(solver, field, iterations, keep)`: 3D, returns `Result3D`.

`backend`: "auto", "gpu" (alias "cuda", the reference's name for its GPU backend). This build has
no CPU backend: "cpu" raises RuntimeError, anything else ValueError. `precision`: "fp64" (complex128
bulk, parity with the FP64 reference) or "fp32" (complex64 bulk with FP64 tables and compensated
arithmetic).
"""

from __future__ import annotations

from typing import Any

import numpy as np

from bwpm_amd import _core
from bwpm_amd.bidirectional import BidirectionalResult
from bwpm_amd.fpwpm import Propagator2D
from bwpm_amd.volume3d import Bidirectional3D, Result3D

CORE_AVAILABLE = True  # the compiled engine is loaded (import of _core fails otherwise)

transmission_matrix = _core.transmission_matrix
reflection_matrix = _core.reflection_matrix


def gpu_available() -> bool:
    """True if an MI355X (any HIP device) is visible."""
    return _core.gpu_device_count() > 0


cuda_available = gpu_available  # the reference's name


def _check_backend(backend: str) -> None:
    if backend in ("auto", "gpu", "cuda"):
        if not gpu_available():
            raise RuntimeError("GPU backend requested but no device is present")
        return
    if backend in ("cpu", "python"):
        raise RuntimeError(f"backend {backend!r} is not part of this build (the engine runs on the GPU)")
    raise ValueError("backend must be 'auto', 'gpu' or 'cuda'")


def run_2d(propagator: Propagator2D, field: Any, iterations: int = 3, keep: bool = False,
           threads: int | None = None, backend: str = "auto", device: int = 0,
           precision: str | None = None) -> BidirectionalResult:
    """2D run (iterations = 0: forward FPWPM only), mapped onto the 3D engine with nx = 1, dx = 1
    (src/bwpm/fast.py:101-162); `threads` is accepted and unused."""
    _check_backend(backend)
    g = propagator.grid
    labels = np.asarray(propagator.labels, dtype=np.int32)[:, :, None]
    nz = labels.shape[0]
    mask = None if propagator.mask is None else np.asarray(propagator.mask)[:, None]
    fl = None if propagator.fluctuation is None else np.asarray(propagator.fluctuation, np.complex128)[:, :, None]
    prec = precision or getattr(propagator, "precision", "fp64")
    solver = _core.GpuSolver(g.ny, 1, g.dy, 1.0, g.dz, g.k0, np.asarray(propagator.indices, dtype=np.complex128),
                             labels, complex(propagator.index_before), mask, device, fl, prec)
    f = np.asarray(field, dtype=np.complex128)
    if f.shape != (g.ny, 2):
        raise ValueError("field must have shape (ny, 2)")
    out = solver.run(f[:, None, :], iterations, list(range(nz + 1)) if keep else [])
    forward = backward = None
    if keep:
        forward = np.stack([out["forward"][p][:, 0] for p in range(nz + 1)])
        if iterations > 0:
            backward = np.stack([out["backward"][p][:, 0] for p in range(nz + 1)])
    residual: list[float] = []
    if iterations > 0:
        k0, dy = g.k0, g.dy
        p_in = _core.power_flux(f[:, None, :], dy, 1.0, k0, propagator.index_before, 1, device)
        last = complex(propagator.indices[int(np.asarray(propagator.labels)[-1][g.ny // 2])])
        for t, r in zip(out["transmitted_iter"], out["reflected_iter"]):
            p_r = _core.power_flux(r, dy, 1.0, k0, propagator.index_before, -1, device)
            p_t = _core.power_flux(t, dy, 1.0, k0, last, 1, device)
            residual.append(1.0 - (p_r + p_t) / p_in if p_in > 0 else float("nan"))
    return BidirectionalResult(out["transmitted"][:, 0], out["reflected"][:, 0], forward, backward, residual)


def run_3d(solver: Bidirectional3D, field: Any, iterations: int = 1, keep: tuple[int, ...] = (),
           threads: int | None = None, backend: str = "auto", device: int = 0,
           precision: str | None = None) -> Result3D:
    """3D run (src/bwpm/fast.py:165-200); `threads` is accepted and unused."""
    _check_backend(backend)
    g = solver.grid
    fl = None if solver.fluctuation is None else np.asarray(solver.fluctuation, dtype=np.complex128)
    prec = precision or getattr(solver, "precision", "fp64")
    core = _core.GpuSolver(g.ny, g.nx, g.dy, g.dx, g.dz, g.k0, np.asarray(solver.indices, dtype=np.complex128),
                           np.asarray(solver.labels, dtype=np.int32), complex(solver.index_before), solver.mask,
                           device, fl, prec)
    out = core.run(np.asarray(field), iterations, list(keep))
    zero = np.zeros_like(out["transmitted"])
    planes = {p: (out["forward"].get(p, zero), out["backward"].get(p, zero)) for p in keep}
    return Result3D(out["transmitted"], out["reflected"], planes, int(out["stored_values"]))


__all__ = ["CORE_AVAILABLE", "cuda_available", "gpu_available", "reflection_matrix", "run_2d", "run_3d",
           "transmission_matrix"]
