"""Compact descriptions of solve inputs that are generated where they are used instead of being
passed as (nz, ny, nx) arrays.

- `Uniform(nz, label)`: a label volume with one label in every pixel of every slab (a homogeneous
  host medium), passed to the library without a label array.
- `FluctuationSynthesis(sigma, corr_mm, seed)`: a real, Gaussian-correlated index fluctuation
  generated on the device. Seeded normal noise is filtered by exp(−|q|²·corr²/4) (periodic in x, y
  and z) and scaled to standard deviation sigma. This is the construction of the reference's test
  volumes (tests/test_fluctuation_3d.py:34-41), with the device's own random stream.
- `Geometry(nz, prims, background, z0)`: a label volume made inside the library by the pixel-centre
  rule of src/bwpm/geometry.py:24-35. Slab j, pixel (n, m) has its centre at
  x = (m − nx//2)·dx, y = (n − ny//2)·dy, z = z0 + (j + ½)·dz (mm) and takes the label of the last
  primitive containing it, else `background`. Primitives: `Above` (z > d + a·x + b·y: planar and
  tilted interfaces), `Layer` (zlo < z ≤ zhi) and `Cylinder` (a z-aligned rod within a layer).
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np


class Uniform(NamedTuple):
    nz: int
    label: int = 0


class FluctuationSynthesis(NamedTuple):
    sigma: float
    corr_mm: float
    seed: int


class Above(NamedTuple):
    """Label where z > d + a·x + b·y."""

    label: int
    d: float
    a: float = 0.0
    b: float = 0.0


class Layer(NamedTuple):
    """Label where zlo < z ≤ zhi."""

    label: int
    zlo: float
    zhi: float


class Cylinder(NamedTuple):
    """Label where sqrt((x − cx)² + (y − cy)²) ≤ r and zlo < z ≤ zhi."""

    label: int
    r: float
    zlo: float
    zhi: float
    cx: float = 0.0
    cy: float = 0.0


class Geometry(NamedTuple):
    nz: int
    prims: tuple = ()
    background: int = 0
    z0: float = 0.0

    def packed(self) -> np.ndarray:
        """(nprims, 10) float64 rows (kind, label, a, b, d, zlo, zhi, cx, cy, r) of the C ABI."""
        rows = []
        for p in self.prims:
            if isinstance(p, Above):
                rows.append((0, p.label, p.a, p.b, p.d, 0, 0, 0, 0, 0))
            elif isinstance(p, Layer):
                rows.append((1, p.label, 0, 0, 0, p.zlo, p.zhi, 0, 0, 0))
            elif isinstance(p, Cylinder):
                rows.append((2, p.label, 0, 0, 0, p.zlo, p.zhi, p.cx, p.cy, p.r))
            else:
                raise ValueError(f"unknown geometry primitive {p!r}")
        return np.ascontiguousarray(np.asarray(rows, dtype=np.float64).reshape(-1, 10))


__all__ = ["Above", "Cylinder", "FluctuationSynthesis", "Geometry", "Layer", "Uniform"]
