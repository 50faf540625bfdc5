"""Pins the analysis oracle (oracle/analysis_oracle.py) on the CPU.

The reference's own analytic tests of `bwpm.pilot_imc3d` (tests/test_pilot_imc3d.py:43-105: the
cos θ law of a plane wave's flux for s and p, scaling with index and amplitude, Fresnel at normal
incidence and at Brewster's angle, the reflectance of s, p and mixed waves, the focus of a known
defocus phase) run against the oracle, plus closed forms the reference does not test: Parseval
for a normal plane wave, R + T = 1 of the Fresnel coefficients, the Strehl ratio 1 of a flat
spectrum at zero offset and of a defocused one at its defocus, and a brute-force on-axis field
(explicit angular-spectrum propagation of the whole plane) for the axial profile.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import analysis_oracle as A

LAMBDA_UM = 0.635
STEP = LAMBDA_UM * 1e-3 / 4
K0 = 2 * np.pi / (LAMBDA_UM * 1e-3)
G = (64, 64, STEP, STEP, K0)  # ny, nx, dy, dx, k0 as the reference's test grid


def mode(mx: int, my: int, ex: complex, ey: complex, ny: int = 64, nx: int = 64) -> np.ndarray:
    ph = np.exp(1j * 2 * np.pi * (mx * np.arange(nx)[None, :] / nx + my * np.arange(ny)[:, None] / ny))
    return np.stack([ex * ph, ey * ph], axis=-1)


def angle(mx: int, my: int, n: float) -> float:
    kp = 2 * np.pi * np.hypot(mx / (64 * STEP), my / (64 * STEP))
    return float(np.arcsin(kp / (n * K0)))


@pytest.mark.parametrize("pol", ["s", "p"])
def test_plane_power_follows_cosine_law(pol: str) -> None:
    n = 1.511
    ref = A.plane_power(mode(0, 0, 1.0, 0.0), *G, n)
    th = angle(6, 0, n)
    f = mode(6, 0, 0.0, 1.0) if pol == "s" else mode(6, 0, np.cos(th), 0.0)
    assert A.plane_power(f, *G, n) / ref == pytest.approx(np.cos(th), rel=1e-12)


def test_plane_power_scaling_and_parseval() -> None:
    f = mode(0, 0, 1.0, 0.0)
    assert A.plane_power(2 * f, *G, 1.0) / A.plane_power(f, *G, 1.0) == pytest.approx(4.0)
    assert A.plane_power(f, *G, 1.5) / A.plane_power(f, *G, 1.0) == pytest.approx(1.5)
    # a normal plane wave |E| = 1 has one mode of amplitude N: flux n·N²
    assert A.plane_power(f, *G, 1.5) == pytest.approx(1.5 * (64 * 64) ** 2, rel=1e-14)
    # evanescent modes carry nothing (beyond the FFT's rounding leakage into other modes)
    kmax = int(64 * STEP * K0 / (2 * np.pi)) + 2
    assert A.plane_power(mode(kmax, kmax, 1.0, 1.0), *G, 1.0) < 1e-20 * A.plane_power(f, *G, 1.0)


def test_fresnel_normal_incidence_brewster_and_energy() -> None:
    n1, n2 = 1.0, 1.511
    rs, rp = A.fresnel_reflectances(n1, n2, np.array([1.0]))
    assert rs[0] == pytest.approx(((n2 - n1) / (n2 + n1)) ** 2, rel=1e-14)
    assert rp[0] == pytest.approx(((n2 - n1) / (n2 + n1)) ** 2, rel=1e-14)
    cos_b = np.cos(np.arctan(n2 / n1))
    assert A.fresnel_reflectances(n1, n2, np.array([cos_b]))[1][0] == pytest.approx(0.0, abs=1e-28)
    assert A.fresnel_reflectances(n1, n2, np.array([1e-9]))[0][0] == pytest.approx(1.0, abs=1e-8)
    # R + T = 1 with T_s = (n2 cos t / n1 cos i)|t_s|², t_s = 2 n1 cos i / (n1 cos i + n2 cos t)
    ci = np.linspace(0.05, 1.0, 40)
    ct = np.sqrt(1 - (n1 / n2) ** 2 * (1 - ci**2))
    ts = 2 * n1 * ci / (n1 * ci + n2 * ct)
    tp = 2 * n1 * ci / (n2 * ci + n1 * ct)
    rs, rp = A.fresnel_reflectances(n1, n2, ci)
    assert np.allclose(rs + n2 * ct / (n1 * ci) * ts**2, 1.0, atol=1e-14)
    assert np.allclose(rp + n2 * ct / (n1 * ci) * tp**2, 1.0, atol=1e-14)
    # total internal reflection from the dense side
    assert np.all(A.fresnel_reflectances(n2, n1, np.array([0.1, 0.3]))[0] == 1.0)


def test_reference_reflectance_weights_by_polarisation() -> None:
    n1, n2 = 1.0, 1.511
    th = angle(10, 0, n1)
    p_wave, s_wave = mode(10, 0, np.cos(th), 0.0), mode(10, 0, 0.0, 1.0)
    rs, rp = A.fresnel_reflectances(n1, n2, np.array([np.cos(th)]))
    assert A.reference_reflectance(s_wave, *G, n1, n2) == pytest.approx(rs[0], rel=1e-12)
    assert A.reference_reflectance(p_wave, *G, n1, n2) == pytest.approx(rp[0], rel=1e-12)
    ps, ss = A.plane_power(p_wave, *G, n1), A.plane_power(s_wave, *G, n1)
    assert A.reference_reflectance(p_wave + s_wave, *G, n1, n2) == pytest.approx(
        (ps * rp[0] + ss * rs[0]) / (ps + ss), rel=1e-12)
    with pytest.raises(ValueError, match="no propagating flux"):
        A.reference_reflectance(np.zeros((64, 64, 2), dtype=complex), *G, n1, n2)


def defocused(distance: float, n: float = 1.0, ny: int = 64, nx: int = 64, cut: float = 0.3) -> np.ndarray:
    kx, ky = A.k_grid(ny, nx, STEP, STEP)
    kp = np.hypot(kx, ky)
    kz = np.sqrt(np.maximum((n * K0) ** 2 - kp**2, 0.0))
    spec = np.where(kp <= cut * K0, np.exp(-1j * kz * distance), 0.0)
    shift = np.exp(-1j * (kx * (nx // 2) * STEP + ky * (ny // 2) * STEP))
    return np.stack([np.fft.ifft2(spec * shift), np.zeros((ny, nx), dtype=complex)], axis=-1)


@pytest.mark.parametrize("distance", [0.012, -0.007, 0.0])
def test_axial_focus_finds_known_defocus(distance: float) -> None:
    best, peak = A.axial_focus(defocused(distance), *G, 1.0, np.arange(-0.03, 0.03, 0.0002))
    assert best == pytest.approx(distance, abs=2e-4)
    assert peak > 0


def test_axial_profile_equals_brute_force_propagation() -> None:
    # on-axis field after distance d: propagate the whole plane by exp(i k_z d) over the
    # propagating modes, transform back, take the centre pixel (x = y = 0) and E_z from k·E = 0
    rng = np.random.default_rng(3)
    ny, nx = 24, 20
    f = rng.normal(size=(ny, nx, 2)) + 1j * rng.normal(size=(ny, nx, 2))
    n = 1.3
    kx, ky = A.k_grid(ny, nx, STEP, STEP)
    c = A.cos_theta(kx, ky, n * K0)
    prop = c > 0
    kz = np.where(prop, n * K0 * c, 1.0)
    a = np.fft.fft2(f, axes=(0, 1))
    for d, want in zip([0.0, 0.003, -0.011], A.axial_profile(f, ny, nx, STEP, STEP, K0, n, [0.0, 0.003, -0.011])):
        ph = np.where(prop, np.exp(1j * kz * d), 0.0)
        ex = np.fft.ifft2(a[..., 0] * ph)[ny // 2, nx // 2]
        ey = np.fft.ifft2(a[..., 1] * ph)[ny // 2, nx // 2]
        ez = np.fft.ifft2(np.where(prop, -(kx * a[..., 0] + ky * a[..., 1]) / kz, 0.0) * ph)[ny // 2, nx // 2]
        assert want == pytest.approx(abs(ex) ** 2 + abs(ey) ** 2 + abs(ez) ** 2, rel=1e-10)


def test_strehl_closed_forms() -> None:
    ny = nx = 64
    kx, ky = A.k_grid(ny, nx, STEP, STEP)
    flat = np.exp(1j * 0.7) * np.ones((ny, nx))
    offs = np.arange(-5e-3, 5e-3, 0.02e-3)
    s, best = A.strehl(flat, ny, nx, STEP, STEP, K0, K0, 0.4, offs)
    assert s == pytest.approx(1.0, rel=1e-12)
    assert abs(best) < 0.5 * 0.02e-3
    kz = np.sqrt(np.maximum(K0**2 - kx**2 - ky**2, 0.0))
    d0 = 1.3e-3
    s, best = A.strehl(np.exp(-1j * kz * d0), ny, nx, STEP, STEP, K0, K0, 0.4, offs)
    assert best == pytest.approx(d0, abs=0.02e-3) and s == pytest.approx(1.0, abs=1e-2)
    # 2D: a flat 1D spectrum focuses at zero offset with Strehl 1
    ky1 = 2 * np.pi * np.fft.fftfreq(128, d=STEP)
    v, z = A.strehl_from_spectrum(ky1[np.abs(ky1) < 0.5 * K0], np.ones(int(np.sum(np.abs(ky1) < 0.5 * K0))), K0,
                                  np.linspace(-1e-3, 1e-3, 101))
    assert v == pytest.approx(1.0, rel=1e-12) and z == pytest.approx(0.0, abs=1e-15)
