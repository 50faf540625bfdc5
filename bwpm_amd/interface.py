"""Transmission and reflection matrices of a plane interface for (E_x, E_y), computed in FP64 on the
GPU through the C ABI.

Same functions, arguments and conventions as the reference's `bwpm.interface`
(src/bwpm/interface.py:23-151): fields ∝ exp(i(k·r − ωt)), k_z with Im k_z ≥ 0, the grazing rule
|ñ²k₀² − k⊥²| ≤ 10⁻¹²·|ñ²k₀²| ⇒ k_z = i·10⁻¹²·k₀, Fresnel Convention A, TE = s, TM = p,
lengths in mm and k in 1/mm.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np

from bwpm_amd import _core

GRAZING_TOLERANCE = 1e-12


def longitudinal_wavenumber(index: complex, k0: float, k_perp_sq: Any) -> np.ndarray:
    """k_z = sqrt(ñ²k₀² − |k⊥|²), Im k_z ≥ 0 (Re k_z ≥ 0 if Im k_z = 0), with the grazing rule."""
    return np.asarray(_core.longitudinal_wavenumber(index, k0, np.asarray(k_perp_sq, dtype=np.float64)))


@dataclass(frozen=True)
class FresnelCoefficients:
    """Convention-A amplitude coefficients and the longitudinal wavenumbers per k⊥."""

    r_s: np.ndarray
    r_p: np.ndarray
    t_s: np.ndarray
    t_p: np.ndarray
    kz_in: np.ndarray
    kz_out: np.ndarray


def fresnel_coefficients(k_x: Any, k_y: Any, index_in: complex, index_out: complex, k0: float) -> FresnelCoefficients:
    """r_s, r_p, t_s, t_p (Convention A) and k_z on both sides at the transverse frequencies."""
    kx, ky = np.broadcast_arrays(np.asarray(k_x, dtype=np.float64), np.asarray(k_y, dtype=np.float64))
    out = _core.fresnel_coefficients(kx, ky, index_in, index_out, k0)
    return FresnelCoefficients(*(np.asarray(out[..., j]) for j in range(6)))


def transmission_matrix(k_x: Any, k_y: Any, index_in: complex, index_out: complex, k0: float) -> np.ndarray:
    """T of wende2023 eqs. (2), (3) without propagation factor, shape (..., 2, 2)."""
    kx, ky = np.broadcast_arrays(np.asarray(k_x, dtype=np.float64), np.asarray(k_y, dtype=np.float64))
    return _core.transmission_matrix(kx, ky, index_in, index_out, k0)


def reflection_matrix(k_x: Any, k_y: Any, index_in: complex, index_out: complex, k0: float) -> np.ndarray:
    """R of wende2023 eqs. (8), (9) without propagation factor, shape (..., 2, 2)."""
    kx, ky = np.broadcast_arrays(np.asarray(k_x, dtype=np.float64), np.asarray(k_y, dtype=np.float64))
    return _core.reflection_matrix(kx, ky, index_in, index_out, k0)


__all__ = ["GRAZING_TOLERANCE", "FresnelCoefficients", "fresnel_coefficients", "longitudinal_wavenumber",
           "reflection_matrix", "transmission_matrix"]
