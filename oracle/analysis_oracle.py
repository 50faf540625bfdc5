"""FP64 CPU oracle of the analysis reductions over planes (the step after the slab-stepping path).

TEST INFRASTRUCTURE ONLY. This module is the checker, never the product: only `tests/` and
`__graft_entry__.smoke()` may import it, and the `bwpm_amd` package never does.

It restates in NumPy the reference's analysis of kept planes:

- `spectrum`, `cos_theta`, `polarisation_split`   src/bwpm/pilot_imc3d.py:20-52
- `plane_power`                                  src/bwpm/pilot_imc3d.py:55-63
- `fresnel_reflectances`                         src/bwpm/pilot_imc3d.py:66-79
- `reference_reflectance`                        src/bwpm/pilot_imc3d.py:82-97
- `axial_profile` / `axial_focus`                src/bwpm/pilot_imc3d.py:100-131
- `strehl` (3D, aperture NA)                     experiments/b3_scaling.py:88-97
- `strehl_from_spectrum` (2D)                    src/bwpm/pilot_imc.py:184-197

Grids are given as (ny, nx, dy, dx, k0) with k0 = 2π/λ in 1/mm. How each function is pinned
(tests/test_analysis_oracle.py): the reference's own analytic tests (tests/test_pilot_imc3d.py:
cos θ law for s and p, scaling with index and amplitude, Fresnel at normal incidence and at the
Brewster angle, reflectance of s and p waves and of their mix, the focus of a known defocus phase),
plus closed forms: Parseval for a normal plane wave (n·N²), Fresnel R + T = 1, a flat spectrum's
Strehl ratio of 1 at zero offset, and the brute-force on-axis field of a small grid propagated by
the angular spectrum for the axial profile.
"""

from __future__ import annotations

import numpy as np

GRAZING = 1e-9  # src/bwpm/pilot_imc3d.py:17


def k_grid(ny: int, nx: int, dy: float, dx: float) -> tuple[np.ndarray, np.ndarray]:
    """(k_x, k_y), shape (ny, nx), numpy order (src/bwpm/volume3d.py:62-67)."""
    kx = 2.0 * np.pi * np.fft.fftfreq(nx, d=dx)
    ky = 2.0 * np.pi * np.fft.fftfreq(ny, d=dy)
    kxx, kyy = np.meshgrid(kx, ky)
    return kxx, kyy


def cos_theta(kx: np.ndarray, ky: np.ndarray, k: float) -> np.ndarray:
    """cos θ of every mode in a medium of wavenumber k; 0 below 1e-9 (pilot_imc3d.py:43-52)."""
    cos_sq = 1.0 - (kx**2 + ky**2) / k**2
    c = np.sqrt(np.maximum(cos_sq, 0.0))
    return np.where(c > GRAZING, c, 0.0)


def polarisation_split(a: np.ndarray, kx: np.ndarray, ky: np.ndarray, cos_t: np.ndarray):
    """|E_s|² and |E_p|² per mode; s along ẑ × k̂, p in the plane of k and ẑ, |E_p| = |E_t·k̂⊥|/cos θ;
    at k⊥ = 0 ŝ = ŷ and k̂⊥ = x̂ (pilot_imc3d.py:26-40)."""
    kp = np.hypot(kx, ky)
    on = kp > 0
    safe = np.where(on, kp, 1.0)
    sx, sy = np.where(on, -ky / safe, 0.0), np.where(on, kx / safe, 1.0)
    px, py = np.where(on, kx / safe, 1.0), np.where(on, ky / safe, 0.0)
    e_s = a[..., 0] * sx + a[..., 1] * sy
    e_p = (a[..., 0] * px + a[..., 1] * py) / np.where(cos_t > GRAZING, cos_t, 1.0)
    return np.abs(e_s) ** 2, np.abs(e_p) ** 2


def plane_power(field, ny, nx, dy, dx, k0, index) -> float:
    """n·Σ cos θ (|E_s|² + |E_p|²) of the unnormalised 2D FFT (pilot_imc3d.py:55-63)."""
    a = np.fft.fft2(np.asarray(field, dtype=np.complex128), axes=(0, 1))
    kx, ky = k_grid(ny, nx, dy, dx)
    c = cos_theta(kx, ky, index * k0)
    s2, p2 = polarisation_split(a, kx, ky, c)
    return float(index * np.sum(c * (s2 + p2)))


def fresnel_reflectances(n1: float, n2: float, cos_i):
    """(R_s, R_p) of n1 → n2, Snell's law, 1 under total internal reflection (pilot_imc3d.py:66-79)."""
    cos_i = np.asarray(cos_i, dtype=np.float64)
    sin_t2 = (n1 / n2) ** 2 * (1.0 - cos_i**2)
    tir = sin_t2 >= 1.0
    ct = np.sqrt(np.where(tir, 0.0, 1.0 - sin_t2))
    with np.errstate(invalid="ignore", divide="ignore"):
        rs = (n1 * cos_i - n2 * ct) / (n1 * cos_i + n2 * ct)
        rp = (n2 * cos_i - n1 * ct) / (n2 * cos_i + n1 * ct)
    return np.where(tir, 1.0, rs**2), np.where(tir, 1.0, rp**2)


def reference_reflectance(field, ny, nx, dy, dx, k0, n1, n2) -> float:
    """Σ cos θ (|E_s|² R_s + |E_p|² R_p) / Σ cos θ (|E_s|² + |E_p|²) in medium n1 (pilot_imc3d.py:82-97)."""
    a = np.fft.fft2(np.asarray(field, dtype=np.complex128), axes=(0, 1))
    kx, ky = k_grid(ny, nx, dy, dx)
    c = cos_theta(kx, ky, n1 * k0)
    s2, p2 = polarisation_split(a, kx, ky, c)
    rs, rp = fresnel_reflectances(n1, n2, c)
    total = float(np.sum(c * (s2 + p2)))
    if total == 0.0:
        raise ValueError("the field carries no propagating flux")
    return float(np.sum(c * (s2 * rs + p2 * rp)) / total)


def axial_profile(field, ny, nx, dy, dx, k0, index, offsets, na=None) -> np.ndarray:
    """(|Σ A_x e^{ik_z d}|² + |Σ A_y …|² + |Σ A_z …|²)/(nx·ny)² per offset d, A referred to the grid
    centre, A_z = −(k_x A_x + k_y A_y)/k_z, propagating modes (within na) only (pilot_imc3d.py:115-131)."""
    a = np.fft.fft2(np.asarray(field, dtype=np.complex128), axes=(0, 1))
    kx, ky = k_grid(ny, nx, dy, dx)
    a = a * np.exp(1j * (kx * (nx // 2) * dx + ky * (ny // 2) * dy))[..., None]
    c = cos_theta(kx, ky, index * k0)
    prop = c > 0
    if na is not None:
        prop &= np.hypot(kx, ky) <= na * k0
    kz = (index * k0 * c)[prop]
    ax, ay = a[..., 0][prop], a[..., 1][prop]
    az = -(kx[prop] * ax + ky[prop] * ay) / kz
    out = []
    for d in np.asarray(offsets, dtype=np.float64):
        ph = np.exp(1j * kz * d)
        out.append(sum(abs(np.sum(v * ph)) ** 2 for v in (ax, ay, az)) / (nx * ny) ** 2)
    return np.asarray(out, dtype=np.float64)


def axial_focus(field, ny, nx, dy, dx, k0, index, offsets, na=None) -> tuple[float, float]:
    """First offset of the maximum and the maximum (pilot_imc3d.py:124-131)."""
    best, peak = 0.0, -1.0
    for d, v in zip(np.asarray(offsets, dtype=np.float64), axial_profile(field, ny, nx, dy, dx, k0, index, offsets, na)):
        if v > peak:
            best, peak = float(d), float(v)
    return best, peak


def strehl_profile(spectrum, ny, nx, dy, dx, k0, k, na, offsets) -> np.ndarray:
    """|Σ A e^{ik_z d}|² / (Σ|A|)² over |k⊥| ≤ na·k0, k_z = sqrt(k² − k⊥²) (b3_scaling.py:88-97)."""
    kx, ky = k_grid(ny, nx, dy, dx)
    support = np.sqrt(kx * kx + ky * ky) <= na * k0
    a = np.asarray(spectrum, dtype=np.complex128)[support]
    kz = np.sqrt(k * k - kx[support] ** 2 - ky[support] ** 2)
    norm = float(np.sum(np.abs(a))) ** 2
    return np.asarray([abs(np.sum(a * np.exp(1j * kz * d))) ** 2 / norm for d in offsets], dtype=np.float64)


def strehl(spectrum, ny, nx, dy, dx, k0, k, na, offsets) -> tuple[float, float]:
    values = strehl_profile(spectrum, ny, nx, dy, dx, k0, k, na, offsets)
    j = int(np.argmax(values))
    return float(values[j]), float(np.asarray(offsets)[j])


def strehl_from_spectrum(ky, spectrum, k, offsets) -> tuple[float, float]:
    """2D: k_z = sqrt(max(k² − k_y², 0)), max over offsets of |Σ A e^{ik_z z}|²/(Σ|A|)² (pilot_imc.py:184-197)."""
    ky = np.asarray(ky, dtype=np.float64)
    kz = np.sqrt(np.maximum(k * k - ky * ky, 0.0))
    spectrum = np.asarray(spectrum, dtype=np.complex128)
    norm = float(np.sum(np.abs(spectrum))) ** 2
    values = [float(np.abs(np.sum(spectrum * np.exp(1j * kz * dz))) ** 2) / norm for dz in offsets]
    j = int(np.argmax(values))
    return values[j], float(np.asarray(offsets)[j])
