"""Analysis of the planes a 3D solve returns, on the device: flux through a plane, the flux-weighted
Fresnel reference reflectance, the axial focus and the Strehl ratio.

Same functions and conventions as the reference's `bwpm.pilot_imc3d` (src/bwpm/pilot_imc3d.py:19-131)
and the Strehl ratio of experiments/b3_scaling.py:88-97: lengths in mm, k in 1/mm, transverse field
(E_x, E_y) of shape (ny, nx, 2) on the centred grid of `Grid3D`. The transforms and every sum over
modes run in FP64 on the GPU (bwpm_plane_power, bwpm_reference_reflectance, bwpm_axial_profile,
bwpm_strehl_profile); only the choice of the maximum over the scanned offsets is made here. A field
can be a NumPy array or a plane left on the device by a solver (`GpuSolver.device_plane`), which
is then read without a copy to the host.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from bwpm_amd import _core
from bwpm_amd.volume3d import Grid3D

NA = 0.4  # numerical aperture of the Strehl evaluation (experiments/b3_scaling.py:33)
STREHL_OFFSETS = np.arange(-5e-3, 5e-3, 0.02e-3)  # experiments/b3_scaling.py:94

_handles: dict[tuple[int, int, float, float, float, int], _core.GpuAnalysis] = {}


def analysis(grid: Grid3D, device: int = 0) -> _core.GpuAnalysis:
    """The device analysis handle of a grid (one per grid and device, kept for reuse)."""
    key = (grid.ny, grid.nx, grid.dy, grid.dx, grid.k0, device)
    h = _handles.get(key)
    if h is None:
        if len(_handles) >= 16:
            _handles.pop(next(iter(_handles))).close()
        h = _handles[key] = _core.GpuAnalysis(grid.ny, grid.nx, grid.dy, grid.dx, grid.k0, device)
    return h


def plane_power(field: Any, grid: Grid3D, index: float) -> float:
    """Flux through the plane, n·Σ cos θ (|E_s|² + |E_p|²) over the propagating modes, in the
    reference's fixed arbitrary units (src/bwpm/pilot_imc3d.py:55-63)."""
    return float(analysis(grid).plane_power(field, index)[0])


def plane_powers(planes: Any, grid: Grid3D, index: float) -> np.ndarray:
    """plane_power of a stack (count, ny, nx, 2), up to four planes per transform launch."""
    return analysis(grid).plane_power(planes, index)


def fresnel_reflectances(n1: float, n2: float, cos_i: Any) -> tuple[np.ndarray, np.ndarray]:
    """Power reflectances (R_s, R_p) of the lossless interface n1 → n2 at incidence cosines cos_i;
    total internal reflection gives 1 (src/bwpm/pilot_imc3d.py:66-79). Host arithmetic on the
    given angles; the per-mode weighting of a field runs on the device in reference_reflectance."""
    cos_i = np.asarray(cos_i, dtype=np.float64)
    sin_t2 = (n1 / n2) ** 2 * (1.0 - cos_i**2)
    tir = sin_t2 >= 1.0
    cos_t = np.sqrt(np.where(tir, 0.0, 1.0 - sin_t2))
    with np.errstate(invalid="ignore", divide="ignore"):
        rs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t)
        rp = (n2 * cos_i - n1 * cos_t) / (n2 * cos_i + n1 * cos_t)
    return np.where(tir, 1.0, rs**2), np.where(tir, 1.0, rp**2)


def reference_reflectance(field: Any, grid: Grid3D, n1: float, n2: float) -> float:
    """Flux-weighted Fresnel reflectance of one interface n1 → n2 for the incident field
    (src/bwpm/pilot_imc3d.py:82-97). ValueError if the field carries no propagating flux."""
    return analysis(grid).reference_reflectance(field, n1, n2)


def axial_profile(field: Any, grid: Grid3D, index: float, offsets: Any, na: float | None = None) -> np.ndarray:
    """On-axis intensity (|E_x|² + |E_y|² + |E_z|²)/(nx·ny)² after each distance of `offsets`."""
    return analysis(grid).axial_profile(field, index, offsets, na)


def axial_focus(field: Any, grid: Grid3D, index: float, offsets: Any, na: float | None = None) -> tuple[float, float]:
    """(offset of the on-axis intensity maximum, the maximum): the first offset whose intensity
    exceeds every earlier one (src/bwpm/pilot_imc3d.py:100-131)."""
    off = np.asarray(offsets, dtype=np.float64).ravel()
    values = axial_profile(field, grid, index, off, na)
    best, peak = 0.0, -1.0
    v = np.where(np.isnan(values), -np.inf, values)
    if v.size:
        j = int(np.argmax(v))
        if v[j] > peak:
            best, peak = float(off[j]), float(v[j])
    return best, peak


def strehl(spectrum: Any, grid: Grid3D, k: float, na: float = NA, offsets: Any = None) -> tuple[float, float]:
    """3D Strehl ratio max_d |Σ A e^(i k_z d)|² / (Σ|A|)² over the modes with |k⊥| ≤ na·k0 of a
    spectrum (ny, nx) referred to the grid centre, and the best offset (mm)
    (experiments/b3_scaling.py:88-97)."""
    off = STREHL_OFFSETS if offsets is None else np.asarray(offsets, dtype=np.float64).ravel()
    values = analysis(grid).strehl_profile(spectrum, k, na, off)
    j = int(np.argmax(values))
    return float(values[j]), float(off[j])


def strehl_from_spectrum(ky: Any, spectrum: Any, k: float, offsets: Any, device: int = 0) -> tuple[float, float]:
    """2D Strehl ratio max_z |Σ A e^(i k_z z)|² / (Σ|A|)² of a spectrum A(k_y) and the best focus
    offset (mm), k_z = sqrt(max(k² − k_y², 0)) (src/bwpm/pilot_imc.py:184-197)."""
    off = np.asarray(offsets, dtype=np.float64).ravel()
    kyv = np.asarray(ky, dtype=np.float64).ravel()
    g = Grid3D(nx=2, ny=2, dx=1.0, dy=1.0, dz=1.0, wavelength_um=1.0)  # the handle supplies the device
    values = analysis(g, device).strehl_modes(kyv, spectrum, k, off)
    j = int(np.argmax(values))
    return float(values[j]), float(off[j])


__all__ = ["NA", "STREHL_OFFSETS", "analysis", "axial_focus", "axial_profile", "fresnel_reflectances", "plane_power",
           "plane_powers", "reference_reflectance", "strehl", "strehl_from_spectrum"]
