"""3D BWPM (x, y, z) with reflection sources kept at interface pixels only, on the GPU.

Same types as the reference's `bwpm.volume3d` (src/bwpm/volume3d.py:34-344): fields (ny, nx, 2)
with components (E_x, E_y) on the centred grid x_m = (m − nx//2)·dx, y_n = (n − ny//2)·dy (mm);
planes p = 0 … nz at z₀ + p·dz, plane p belongs to slab p − 1; optional weak index fluctuations δn
scatter with the vector Born approximation (forward increment and R1 in the forward pass, R2 and R3
in the backward pass).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np


@dataclass(frozen=True)
class Grid3D:
    """Transverse grid (ny × nx) and step; lengths in mm, wavelength in µm."""

    nx: int
    ny: int
    dx: float
    dy: float
    dz: float
    wavelength_um: float

    def __post_init__(self) -> None:
        if min(self.nx, self.ny) < 2 or min(self.dx, self.dy, self.dz, self.wavelength_um) <= 0:
            raise ValueError("nx, ny >= 2 and positive steps and wavelength are required")

    @property
    def k0(self) -> float:
        return float(2.0 * np.pi / (self.wavelength_um * 1e-3))

    @property
    def x(self) -> np.ndarray:
        return np.asarray((np.arange(self.nx) - self.nx // 2) * self.dx, dtype=np.float64)

    @property
    def y(self) -> np.ndarray:
        return np.asarray((np.arange(self.ny) - self.ny // 2) * self.dy, dtype=np.float64)

    def k_grid(self) -> tuple[np.ndarray, np.ndarray]:
        """(k_x, k_y) of the 2D FFT, shape (ny, nx) each, 1/mm (numpy order)."""
        kx = 2.0 * np.pi * np.fft.fftfreq(self.nx, d=self.dx)
        ky = 2.0 * np.pi * np.fft.fftfreq(self.ny, d=self.dy)
        kxx, kyy = np.meshgrid(kx, ky)
        return np.asarray(kxx), np.asarray(kyy)


def absorbing_mask_2d(grid: Grid3D, width: float, order: int = 8, edge: float = 1e-6) -> np.ndarray:
    """Product of 1D super-Gaussian edge absorbers in x and y, shape (ny, nx)."""
    if width < 0 or order < 1 or not 0 < edge < 1:
        raise ValueError("width >= 0, order >= 1 and 0 < edge < 1 are required")

    def one(c: np.ndarray) -> np.ndarray:
        a = np.abs(c)
        if width == 0:
            return np.ones_like(a)
        u = np.clip((a - (float(np.max(a)) - width)) / width, 0.0, 1.0)
        return np.asarray(edge ** (u ** order))

    return np.asarray(one(grid.y)[:, None] * one(grid.x)[None, :])


@dataclass(frozen=True)
class Result3D:
    """transmitted (ny, nx, 2) at plane nz, reflected at plane 0, requested planes
    {p: (forward, backward)}, and the number of stored reflection source vectors."""

    transmitted: np.ndarray
    reflected: np.ndarray
    planes: dict[int, tuple[np.ndarray, np.ndarray]]
    stored_values: int


@dataclass
class Bidirectional3D:
    """Forward FPWPM and BWPM in 3D on the GPU.

    `sparse` and `workers` of the reference are accepted for signature compatibility: reflection
    sources are always sparse here and the device does the parallel work. `precision` selects the
    bulk precision ("fp64" as the reference, or "fp32")."""

    grid: Grid3D
    indices: list[complex]
    labels: Any
    index_before: complex
    mask: Any = None
    sparse: bool = True
    workers: int | None = None
    fluctuation: Any = None
    precision: str = "fp64"
    device: int = 0
    _hosts: list[complex | None] = field(default_factory=list, repr=False)

    def __post_init__(self) -> None:
        lab = np.asarray(self.labels)
        g = self.grid
        if lab.ndim != 3 or lab.shape[1:] != (g.ny, g.nx):
            raise ValueError("labels must have shape (nz, ny, nx)")
        if lab.min() < 0 or lab.max() >= len(self.indices):
            raise ValueError("labels must index into `indices`")

    def run(self, field_in: Any, iterations: int = 1, keep: tuple[int, ...] = ()) -> Result3D:
        """Forward-backward iterations; `iterations = 0` runs the forward FPWPM only."""
        from bwpm_amd import fast

        return fast.run_3d(self, field_in, iterations, keep, device=self.device, precision=self.precision)


def memory_estimate(nx: int, ny: int, interface_pixels: int, pairs: int, kinds: int = 4, nz: int = 0,
                    label_bytes: int = 1, precision: str = "fp64") -> int:
    """Device bytes of one solve: operator tables (hi + lo, three complex per k and kind), working
    planes (input, row spectrum, products, field, stage), sparse sources (two value sets, pixel index
    and pair byte per interface pixel). Labels stay on the host; the per-pixel pair map of a
    non-uniform boundary costs one byte per pixel, counted through `nz`·`label_bytes` for the
    boundaries that have one (pass nz = 0 for planar stacks)."""
    c = 16 if precision == "fp64" else 8
    plane = nx * ny
    return int(kinds * pairs * plane * 6 * c + 8 * plane * 2 * c + 2 * interface_pixels * (2 * 2 * c + 5)
               + nz * plane * label_bytes)


__all__ = ["Bidirectional3D", "Grid3D", "Result3D", "absorbing_mask_2d", "memory_estimate"]
