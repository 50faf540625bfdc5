"""The flux entry point without device allocation (bwpm_power_flux_noalloc) against the allocating
one (bwpm_power_flux), through the C ABI: the same bits on two calls in a row, the second on a
smaller grid, and a clear error beyond its size limit."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ARGS = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
        ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]


def _entry(name: str):
    from bwpm_amd import _core

    assert _core.gpu_device_count() > 0, "no GPU visible"
    f = getattr(_core._lib, name)
    f.argtypes = ARGS
    f.restype = ctypes.c_int
    return f


def _flux(name: str, field: np.ndarray, dy: float, dx: float, k0: float, index: complex, direction: int) -> tuple[int, float]:
    f = np.ascontiguousarray(field, dtype=np.complex128)
    out = ctypes.c_double(0.0)
    st = _entry(name)(f.ctypes.data, f.shape[0], f.shape[1], dy, dx, k0, index.real, index.imag, direction, 0,
                      ctypes.byref(out))
    return st, float(out.value)


def _field(rng: np.random.Generator, ny: int, nx: int, dy: float, dx: float, k0: float) -> np.ndarray:
    f = rng.normal(size=(ny, nx, 2)) + 1j * rng.normal(size=(ny, nx, 2))
    kx = 2 * np.pi * np.fft.fftfreq(nx, dx)[None, :]
    ky = 2 * np.pi * np.fft.fftfreq(ny, dy)[:, None]
    sp = np.fft.fft2(f, axes=(0, 1))
    sp[np.sqrt(kx ** 2 + ky ** 2) >= 1.2 * k0] = 0.0
    return np.fft.ifft2(sp, axes=(0, 1))


@pytest.mark.parametrize("direction", [1, -1])
def test_noalloc_flux_equals_the_allocating_flux(direction: int) -> None:
    rng = np.random.default_rng(17)
    k0 = 2 * np.pi / 0.8e-3
    # a full-size grid, then a smaller one (the scratch keeps the first call's contents), then a 2D
    # grid (one column) and a grid whose line lengths need the generic radices
    for ny, nx, index in ((512, 512, 1.5 + 0j), (96, 64, 1.0 + 0j), (4096, 1, 1.33 + 0.01j), (120, 80, 1.2 + 0j)):
        dy, dx = 0.8e-3 / 6, 0.8e-3 / 5
        f = _field(rng, ny, nx, dy, dx, k0)
        s1, want = _flux("bwpm_power_flux", f, dy, dx, k0, index, direction)
        s2, got = _flux("bwpm_power_flux_noalloc", f, dy, dx, k0, index, direction)
        print(f"noalloc flux {ny}x{nx} direction {direction}: {got!r} vs {want!r}", flush=True)
        assert s1 == 0 and s2 == 0
        assert got == want


def test_noalloc_flux_rejects_grids_beyond_its_limit() -> None:
    from bwpm_amd import _core

    f = np.zeros((1024, 512, 2), dtype=np.complex128)
    st, _ = _flux("bwpm_power_flux_noalloc", f, 1e-4, 1e-4, 1e4, 1.0 + 0j, 1)
    assert st == _core.EINVAL
    assert b"no-allocation flux limit" in _core._lib.bwpm_last_error()
