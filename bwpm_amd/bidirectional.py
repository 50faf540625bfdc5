"""2D BWPM: forward/backward iterations with reflections (on the GPU).

Same types as the reference's `bwpm.bidirectional` (src/bwpm/bidirectional.py:44-203): planes
p = 0 … nz, plane p belongs to slab p − 1; one iteration is a forward pass (input plus the sources
of the last backward pass) and a backward pass (sources of this forward pass); the energy residual
ε = 1 − (P_R + P_T)/P_in is reported after every iteration.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np

from bwpm_amd.fpwpm import Propagator2D


@dataclass(frozen=True)
class BidirectionalResult:
    """transmitted (ny, 2) at plane nz; reflected (ny, 2) at plane 0; all planes (nz + 1, ny, 2)
    if requested; energy residual per iteration."""

    transmitted: np.ndarray
    reflected: np.ndarray
    forward: np.ndarray | None
    backward: np.ndarray | None
    energy_residual: list[float]


class Bidirectional2D:
    """BWPM on top of the forward propagator (grid, indices, labels, mask, δn)."""

    def __init__(self, propagator: Propagator2D) -> None:
        self.p = propagator

    def run(self, field_in: Any, iterations: int = 3, keep: bool = False) -> BidirectionalResult:
        """Run `iterations` forward-backward cycles for the input field (ny, 2) at plane 0."""
        from bwpm_amd import fast

        if iterations < 1:
            raise ValueError("iterations must be >= 1")
        f = np.asarray(field_in, dtype=np.complex128)
        if f.shape != (self.p.grid.ny, 2):
            raise ValueError("field must have shape (ny, 2)")
        return fast.run_2d(self.p, f, iterations, keep, device=self.p.device, precision=self.p.precision)


__all__ = ["Bidirectional2D", "BidirectionalResult"]
