"""2D FPWPM (y, z; infinite in x): grid, absorbing mask, forward propagator, E_z and power flux.

Same types and semantics as the reference's `bwpm.fpwpm` (src/bwpm/fpwpm.py:30-216): the volume is
`nz` slabs of thickness dz sampled at `ny` points (centred grid y_n = (n − ny//2)·dy, mm); every
pixel carries a material label into a list of complex indices; a field (ny, 2) = (E_x, E_y) sits at
the right edge of its slab. Propagation runs on the GPU (the 2D problem is the 3D engine with
nx = 1, dx = 1, as src/bwpm/fast.py:120-142 maps it).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np

from bwpm_amd import _core


@dataclass(frozen=True)
class Grid2D:
    """Transverse grid and step: ny samples of dy (mm), slab thickness dz (mm), wavelength (µm)."""

    ny: int
    dy: float
    dz: float
    wavelength_um: float

    def __post_init__(self) -> None:
        if self.ny < 2 or self.dy <= 0 or self.dz <= 0 or self.wavelength_um <= 0:
            raise ValueError("ny >= 2 and positive dy, dz, wavelength_um are required")

    @property
    def k0(self) -> float:
        """Vacuum wavenumber 2π/λ in 1/mm."""
        return float(2.0 * np.pi / (self.wavelength_um * 1e-3))

    @property
    def y(self) -> np.ndarray:
        return np.asarray((np.arange(self.ny) - self.ny // 2) * self.dy, dtype=np.float64)

    @property
    def ky(self) -> np.ndarray:
        """Transverse frequencies of the FFT in 1/mm (numpy order)."""
        return np.asarray(2.0 * np.pi * np.fft.fftfreq(self.ny, d=self.dy), dtype=np.float64)


def absorbing_mask(grid: Grid2D, width: float, order: int = 4, edge: float = 1e-6) -> np.ndarray:
    """Super-Gaussian edge absorber m(y): 1 inside, edge^(u^order) over the boundary zone of `width`."""
    if width < 0 or order < 1 or not 0 < edge < 1:
        raise ValueError("width >= 0, order >= 1 and 0 < edge < 1 are required")
    y = np.abs(grid.y)
    if width == 0:
        return np.ones_like(y)
    u = np.clip((y - (float(np.max(y)) - width)) / width, 0.0, 1.0)
    return np.asarray(edge ** (u ** order), dtype=np.float64)


@dataclass
class Propagator2D:
    """Forward FPWPM through a labelled volume (labels (nz, ny), optional mask (ny,) and δn (nz, ny)).

    `precision` selects the GPU bulk precision ("fp64" as the reference, or "fp32")."""

    grid: Grid2D
    indices: list[complex]
    labels: Any
    index_before: complex
    mask: Any = None
    fluctuation: Any = None
    precision: str = "fp64"
    device: int = 0

    def __post_init__(self) -> None:
        labels = np.asarray(self.labels)
        if labels.ndim != 2 or labels.shape[1] != self.grid.ny:
            raise ValueError("labels must have shape (nz, ny)")
        if labels.min() < 0 or labels.max() >= len(self.indices):
            raise ValueError("labels must index into `indices`")

    def step(self, field_in: Any, labels_before: Any, labels_after: Any, delta_n: Any = None,
             host: complex | None = None) -> np.ndarray:
        """One step z → z + dz from a plane with `labels_before` (−1 = index_before) into a slab
        with `labels_after`, with the Born increment of `delta_n` (src/bwpm/fpwpm.py:134-161)."""
        g = self.grid
        f = np.asarray(field_in, dtype=np.complex128)
        if f.shape != (g.ny, 2):
            raise ValueError("field must have shape (ny, 2)")
        fl = None
        if delta_n is not None and host is not None:
            fl = np.asarray(delta_n, dtype=np.complex128).reshape(1, g.ny, 1)
        solver = _core.GpuSolver(g.ny, 1, g.dy, 1.0, g.dz, g.k0, np.asarray(self.indices, dtype=np.complex128),
                                 np.asarray(labels_after).reshape(1, g.ny, 1), complex(self.index_before),
                                 None if self.mask is None else np.asarray(self.mask).reshape(g.ny, 1),
                                 self.device, fl, self.precision,
                                 labels_before=np.asarray(labels_before).reshape(g.ny, 1))
        return solver.run(f[:, None, :], 0)["transmitted"][:, 0]

    def run(self, field_in: Any, keep: bool = False) -> np.ndarray:
        """Field at the right edge of the last slab (ny, 2), or all planes (nz + 1, ny, 2)."""
        from bwpm_amd import fast

        f = np.asarray(field_in, dtype=np.complex128)
        if f.shape != (self.grid.ny, 2):
            raise ValueError("field must have shape (ny, 2)")
        res = fast.run_2d(self, f, iterations=0, keep=keep, device=self.device, precision=self.precision)
        return res.forward if keep else res.transmitted


def longitudinal_component(field_xy: Any, grid: Grid2D, index: complex, direction: int = 1) -> np.ndarray:
    """E_z from the transverse components in a homogeneous medium, wende2023 eq. (5)."""
    spectrum = np.fft.fft(np.asarray(field_xy), axis=0)
    kz = _core.longitudinal_wavenumber(index, grid.k0, grid.ky ** 2)
    ez = np.where(kz != 0, -direction * grid.ky * spectrum[:, 1] / kz, 0.0)
    return np.asarray(np.fft.ifft(ez))


def power_flux(field_xy: Any, grid: Grid2D, index: complex, direction: int = 1) -> float:
    """Power through the plane per unit length in x (arbitrary units), E_z from eq. (5); evanescent
    waves carry none. Device FP64 reduction."""
    f = np.asarray(field_xy, dtype=np.complex128).reshape(grid.ny, 1, 2)
    return _core.power_flux(f, grid.dy, 1.0, grid.k0, index, direction)


__all__ = ["Grid2D", "Propagator2D", "absorbing_mask", "longitudinal_component", "power_flux"]
