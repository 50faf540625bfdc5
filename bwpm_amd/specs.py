"""Compact descriptions of solve inputs that are generated where they are used instead of being
passed as (nz, ny, nx) arrays.

- `Uniform(nz, label)`: a label volume with one label in every pixel of every slab (a homogeneous
  host medium), passed to the library without a label array.
- `FluctuationSynthesis(sigma, corr_mm, seed)`: a real, Gaussian-correlated index fluctuation
  generated on the device. Seeded normal noise is filtered by exp(−|q|²·corr²/4) (periodic in x, y
  and z) and scaled to standard deviation sigma. This is the construction of the reference's test
  volumes (tests/test_fluctuation_3d.py:34-41), with the device's own random stream.
"""

from __future__ import annotations

from typing import NamedTuple


class Uniform(NamedTuple):
    nz: int
    label: int = 0


class FluctuationSynthesis(NamedTuple):
    sigma: float
    corr_mm: float
    seed: int


__all__ = ["FluctuationSynthesis", "Uniform"]
