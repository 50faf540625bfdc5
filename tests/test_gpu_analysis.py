"""Analysis reductions on the device (bwpm_plane_power, bwpm_reference_reflectance,
bwpm_axial_profile, bwpm_strehl_profile, bwpm_strehl_modes, bwpm_solver_plane) against the CPU
oracle of src/bwpm/pilot_imc3d.py, experiments/b3_scaling.py and src/bwpm/pilot_imc.py.

FP64 throughout with a different summation order than NumPy's pairwise sums: flux, reflectance
and scan values agree to 1e-12 relative (scan values relative to the largest of the profile), and
the chosen maximum is the same offset. The reference's own analytic cases run through the device
path too. Planes left on the device by a solver give the same numbers, bit for bit, as the same
planes copied to the host (FP64 planes are used as they are, FP32 planes are widened exactly).
"""

from __future__ import annotations

import numpy as np
import pytest

from bwpm_amd import _core, pilot_imc3d as P, workloads
from bwpm_amd.volume3d import Grid3D
from oracle import analysis_oracle as A

pytestmark = pytest.mark.gpu

LAMBDA_UM = 0.635
STEP = LAMBDA_UM * 1e-3 / 4
GRID = Grid3D(nx=64, ny=64, dx=STEP, dy=STEP, dz=STEP, wavelength_um=LAMBDA_UM)


def _g(grid: Grid3D) -> tuple:
    return grid.ny, grid.nx, grid.dy, grid.dx, grid.k0


def _field(grid: Grid3D, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    # a smooth beam plus noise, so that propagating and evanescent modes both carry weight
    x, y = grid.x[None, :], grid.y[:, None]
    w = 0.2 * min(grid.nx * grid.dx, grid.ny * grid.dy)
    beam = np.exp(-(x**2 + y**2) / w**2 + 1j * 3e3 * x)
    f = np.stack([beam, 0.3j * beam], axis=-1)
    return f + 0.05 * (rng.normal(size=f.shape) + 1j * rng.normal(size=f.shape))


GRIDS = [GRID, Grid3D(nx=80, ny=96, dx=STEP, dy=1.2 * STEP, dz=STEP, wavelength_um=LAMBDA_UM),
         Grid3D(nx=256, ny=128, dx=0.8 * STEP, dy=STEP, dz=STEP, wavelength_um=1.55)]


@pytest.mark.parametrize("gi", range(len(GRIDS)))
def test_flux_and_reflectance_equal_the_oracle(gi: int) -> None:
    grid = GRIDS[gi]
    f = _field(grid, gi)
    for n in (1.0, 1.46):
        got = P.plane_power(f, grid, n)
        assert got == pytest.approx(A.plane_power(f, *_g(grid), n), rel=1e-12)
    for n1, n2 in ((1.0, 1.511), (1.5, 1.0), (1.0, 1.0)):
        want = A.reference_reflectance(f, *_g(grid), n1, n2)
        assert P.reference_reflectance(f, grid, n1, n2) == pytest.approx(want, rel=1e-12, abs=1e-300)
    stack = np.stack([_field(grid, 10 + i) for i in range(6)])
    many = P.plane_powers(stack, grid, 1.2)
    for i in range(6):
        assert many[i] == pytest.approx(A.plane_power(stack[i], *_g(grid), 1.2), rel=1e-12)


@pytest.mark.parametrize("na", [None, 0.3])
@pytest.mark.parametrize("gi", range(len(GRIDS)))
def test_axial_profile_equals_the_oracle(gi: int, na: float | None) -> None:
    grid = GRIDS[gi]
    f = _field(grid, 20 + gi)
    offs = np.linspace(-0.02, 0.02, 161)
    got = P.axial_profile(f, grid, 1.1, offs, na)
    want = A.axial_profile(f, *_g(grid), 1.1, offs, na)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(want)
    best, peak = P.axial_focus(f, grid, 1.1, offs, na)
    wbest, wpeak = A.axial_focus(f, *_g(grid), 1.1, offs, na)
    assert best == wbest and peak == pytest.approx(wpeak, rel=1e-12)


def test_strehl_profiles_equal_the_oracle() -> None:
    grid = GRIDS[2]
    rng = np.random.default_rng(5)
    kx, ky = grid.k_grid()
    kz = np.sqrt(np.maximum(grid.k0**2 - kx**2 - ky**2, 0.0))
    spec = np.exp(-1j * kz * 1.1e-3) * (1 + 0.2 * rng.normal(size=kz.shape))
    offs = P.STREHL_OFFSETS
    got = P.analysis(grid).strehl_profile(spec, grid.k0 * 1.2, 0.4, offs)
    want = A.strehl_profile(spec, *_g(grid), grid.k0 * 1.2, 0.4, offs)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(want)
    v, best = P.strehl(spec, grid, grid.k0 * 1.2)
    wv, wbest = A.strehl(spec, *_g(grid), grid.k0 * 1.2, 0.4, offs)
    assert best == wbest and v == pytest.approx(wv, rel=1e-12)
    ky1 = 2 * np.pi * np.fft.fftfreq(300, d=STEP)
    sp1 = np.exp(1j * rng.uniform(0, 0.3, 300)) * np.exp(-(ky1 / (0.3 * grid.k0)) ** 2)
    z = np.linspace(-2e-3, 2e-3, 401)
    v, best = P.strehl_from_spectrum(ky1, sp1, grid.k0, z)
    wv, wbest = A.strehl_from_spectrum(ky1, sp1, grid.k0, z)
    assert best == wbest and v == pytest.approx(wv, rel=1e-12)


# ---- the reference's analytic cases (tests/test_pilot_imc3d.py:43-105) through the device ----

def _mode(mx: int, my: int, ex: complex, ey: complex) -> np.ndarray:
    ph = np.exp(1j * 2 * np.pi * (mx * np.arange(64)[None, :] / 64 + my * np.arange(64)[:, None] / 64))
    return np.stack([ex * ph, ey * ph], axis=-1)


def _angle(mx: int, n: float) -> float:
    return float(np.arcsin(2 * np.pi * mx / (64 * STEP) / (n * GRID.k0)))


@pytest.mark.parametrize("pol", ["s", "p"])
def test_plane_power_follows_cosine_law(pol: str) -> None:
    n = 1.511
    th = _angle(6, n)
    f = _mode(6, 0, 0.0, 1.0) if pol == "s" else _mode(6, 0, np.cos(th), 0.0)
    assert P.plane_power(f, GRID, n) / P.plane_power(_mode(0, 0, 1.0, 0.0), GRID, n) == pytest.approx(np.cos(th),
                                                                                                      rel=1e-12)
    assert P.plane_power(_mode(0, 0, 1.0, 0.0), GRID, 1.5) == pytest.approx(1.5 * 4096**2, rel=1e-14)


def test_reference_reflectance_of_s_p_and_mixed_waves() -> None:
    n1, n2 = 1.0, 1.511
    th = _angle(10, n1)
    p_wave, s_wave = _mode(10, 0, np.cos(th), 0.0), _mode(10, 0, 0.0, 1.0)
    rs, rp = P.fresnel_reflectances(n1, n2, np.array([np.cos(th)]))
    assert P.reference_reflectance(s_wave, GRID, n1, n2) == pytest.approx(rs[0], rel=1e-12)
    assert P.reference_reflectance(p_wave, GRID, n1, n2) == pytest.approx(rp[0], rel=1e-12)
    ps, ss = P.plane_power(p_wave, GRID, n1), P.plane_power(s_wave, GRID, n1)
    assert P.reference_reflectance(p_wave + s_wave, GRID, n1, n2) == pytest.approx(
        (ps * rp[0] + ss * rs[0]) / (ps + ss), rel=1e-12)
    with pytest.raises(ValueError, match="no propagating flux"):
        P.reference_reflectance(np.zeros((64, 64, 2), dtype=complex), GRID, n1, n2)


@pytest.mark.parametrize("distance", [0.012, -0.007, 0.0])
def test_axial_focus_finds_known_defocus(distance: float) -> None:
    kx, ky = GRID.k_grid()
    kp = np.hypot(kx, ky)
    kz = np.sqrt(np.maximum(GRID.k0**2 - kp**2, 0.0))
    spec = np.where(kp <= 0.3 * GRID.k0, np.exp(-1j * kz * distance), 0.0)
    shift = np.exp(-1j * (kx * 32 * GRID.dx + ky * 32 * GRID.dy))
    f = np.stack([np.fft.ifft2(spec * shift), np.zeros((64, 64), dtype=complex)], axis=-1)
    best, peak = P.axial_focus(f, GRID, 1.0, np.arange(-0.03, 0.03, 0.0002))
    assert best == pytest.approx(distance, abs=2e-4) and peak > 0


# ---- planes left on the device by a solver ---------------------------------------------------

@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_device_planes_equal_host_copies(precision: str) -> None:
    p = workloads.tilted_3d(n=64, nz=24, iterations=1)
    grid = Grid3D(nx=p.nx, ny=p.ny, dx=p.dx, dy=p.dy, dz=p.dz, wavelength_um=p.wavelength_um)
    s = _core.GpuSolver(*p.solver_args(), fluctuation=p.fluctuation, precision=precision)
    try:
        res = s.run(p.field, 1, [0, 12])
        cases = [(s.device_plane("transmitted"), res["transmitted"]), (s.device_plane("reflected"), res["reflected"]),
                 (s.device_plane("forward", 12), res["forward"][12]), (s.device_plane("backward", 0), res["backward"][0])]
        offs = np.linspace(-0.01, 0.01, 41)
        for dev, host in cases:
            assert P.plane_power(dev, grid, 1.2) == P.plane_power(host, grid, 1.2)
            assert P.reference_reflectance(dev, grid, 1.0, 1.46) == P.reference_reflectance(host, grid, 1.0, 1.46)
            assert np.array_equal(P.axial_profile(dev, grid, 1.0, offs), P.axial_profile(host, grid, 1.0, offs))
        with pytest.raises(ValueError, match="not kept"):
            s.device_plane("forward", 5)
    finally:
        s.close()


def test_argument_errors() -> None:
    f = _mode(0, 0, 1.0, 0.0)
    with pytest.raises(ValueError):
        P.axial_profile(f, GRID, 1.0, [])
    with pytest.raises(ValueError):
        P.plane_power(f[:32], GRID, 1.0)
    with pytest.raises(ValueError):
        P.plane_power(f, GRID, -1.0)
    with pytest.raises(ValueError, match="positive"):
        P.analysis(GRID).strehl_profile(np.ones((64, 64)), GRID.k0, 0.0, [0.0])
