"""Plugin adapter over the MI355X engine, with the call surface of the reference's `bwpm.adapter`
(src/bwpm/adapter.py:38-310): `describe`, `prepare`, `apply`, `estimate`, `port_power`,
`front_port`, `back_port`, `plane_wave` and the types `Port`, `PortField`, `Structure`, `Handle`.

The names, the dataclass fields and the error messages are the reference's, because callers build
these types and match these messages. Everything behind them is this package's:

- `describe()` is assembled from the engine's own tables (precision modes, determinism, backends);
- `prepare()` hashes a canonical JSON header of the scalar parameters, then each array with its
  dtype and shape, under an engine tag;
- solvers are kept per (wavelength, precision) on the handle;
- `apply()` solves on the GPU through `fast.run_3d`;
- `estimate()` prices a solve with the update rate measured on one MI355X and the engine's memory
  model;
- `port_power()` is the device FP64 flux reduction (`bwpm_power_flux`);
- `plane_wave()` is the outer product of two axis phase vectors.
"""

from __future__ import annotations

import hashlib
import json
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from bwpm_amd import __version__, _core, fast
from bwpm_amd.volume3d import Bidirectional3D, Grid3D, absorbing_mask_2d, memory_estimate

# seconds per grid point and plane pass of one solve alone on one MI355X (512² plain slab steps,
# FP64, one solve per launch; bench.py `checks.plain_one_solve_per_launch` = 1.16e10 updates/s);
# the batched lanes of `bwpm_amd.serve` reach several times this rate
SECONDS_PER_POINT_PASS = {"gpu": 8.6e-11}


@dataclass(frozen=True)
class Port:
    """A port plane: position z (mm), direction ±1, medium index, wavelength (µm), centred grid."""

    z: float
    direction: int
    index: complex
    wavelength_um: float
    nx: int
    ny: int
    dx: float
    dy: float

    def __post_init__(self) -> None:
        if self.direction not in (1, -1):
            raise ValueError("direction must be +1 or -1")


@dataclass(frozen=True)
class PortField:
    """Transverse field (E_x, E_y), shape (ny, nx, 2), at a port."""

    port: Port
    field: Any

    def __post_init__(self) -> None:
        if np.shape(self.field) != (self.port.ny, self.port.nx, 2):
            raise ValueError("field must have shape (ny, nx, 2) of the port grid")


@dataclass(frozen=True)
class Structure:
    """Index volume between z0 and z0 + nz·dz (labels (nz, ny, nx) into `indices`)."""

    nx: int
    ny: int
    dx: float
    dy: float
    dz: float
    indices: tuple[complex, ...]
    labels: Any
    index_before: complex
    z0: float = 0.0
    mask_width: float = 0.0


@dataclass
class Handle:
    """Prepared structure; `content_hash` identifies structure, wavelengths and engine version."""

    structure: Structure
    wavelengths_um: tuple[float, ...]
    content_hash: str
    _solvers: dict[tuple[float, str], Bidirectional3D] = field(default_factory=dict, repr=False)


# ---- self-description ------------------------------------------------------------------------

_METHOD = ("fast bidirectional vector wave propagation (Wende et al., Opt. Express 31, 28874, 2023): "
           "angular-spectrum slab steps with 2x2 Fresnel interface operators, iterated forward and "
           "backward; hand-written HIP kernels on one MI355X")

_CONVENTIONS = {
    "time": "harmonic factor exp(-i omega t); a plane wave is exp(i(k.r - omega t))",
    "fresnel": "Convention A of Byrnes, arXiv:1603.02720v5, eq. 6",
    "polarisation": "TE is s, TM is p; ports carry the transverse components (E_x, E_y)",
    "units": {"length": "mm", "wavelength": "um", "k": "1/mm"},
    "grid": "pixel m at x = (m - nx//2) dx (centred); spectra refer to the origin (0, 0)",
    "kz": "branch Im k_z >= 0; |n^2 k0^2 - kperp^2| <= 1e-12 |n^2 k0^2| counts as grazing",
}

_PORTS = (
    {"name": "front", "plane": "z0", "in": "+z", "out": "-z (reflected)"},
    {"name": "back", "plane": "z0 + nz dz", "out": "+z (transmitted)"},
)

_VALIDITY = (
    "reflection amplitudes are quantitative at interfaces normal to z; at tilted surfaces only "
    "their direction and phase are",
    "transmitted power through tilted surfaces holds to a few percent",
    "fields trapped by total internal reflection do not settle over iterations: bound the count",
    "absorbing edge zones need at least 20 wavelengths (super-Gaussian of order 8)",
)

_PRECISION = {
    "fp64": "complex128 planes and arithmetic (the reference's)",
    "mixed": "complex64 streaming planes, FP64 arithmetic and operator tables",
    "bfp": "block-floating-point streaming planes of complex64 size (28-bit mantissas), FP64 arithmetic",
    "fp32": "complex64 planes and arithmetic, FP64 operator tables (compensated)",
}


def _backends() -> list[str]:
    return ["gpu"] if fast.gpu_available() else []


def describe() -> dict[str, Any]:
    """Self-description for a planner: method, conventions, ports, properties, validity."""
    props = {"linear": True, "reciprocal": True, "operator": False}
    return {
        "name": "BWPM",
        "version": __version__,
        "method": _METHOD,
        "conventions": {k: (dict(v) if isinstance(v, dict) else v) for k, v in _CONVENTIONS.items()},
        "ports": [dict(p) for p in _PORTS],
        "representation": "gridded transverse field",
        **props,
        "deterministic": {"gpu": "bitwise (fixed reduction order; batched and served solves equal solves alone)"},
        "precision": dict(_PRECISION),
        "backends": _backends(),
        "validity": list(_VALIDITY),
    }


# ---- preparation ------------------------------------------------------------------------------

def _content_hash(s: Structure, labels: np.ndarray, wavelengths: np.ndarray) -> str:
    def cplx(z: complex) -> list[float]:
        z = complex(z)
        return [z.real, z.imag]

    header = {"engine": "bwpm_amd", "version": __version__,
              "grid": [int(s.nx), int(s.ny), float(s.dx), float(s.dy), float(s.dz)],
              "index_before": cplx(s.index_before), "z0": float(s.z0), "mask_width": float(s.mask_width)}
    h = hashlib.sha256(json.dumps(header, sort_keys=True).encode())
    for arr in (np.asarray(s.indices, dtype=np.complex128), labels, wavelengths):
        h.update(f"{arr.dtype.str}{arr.shape}".encode())
        h.update(np.ascontiguousarray(arr).tobytes())
    return h.hexdigest()


def prepare(structure: Structure, wavelengths_um: list[float] | tuple[float, ...]) -> Handle:
    """Check the structure and return a handle with a content hash (nothing is computed)."""
    labels = np.asarray(structure.labels)
    if labels.ndim != 3 or labels.shape[1:] != (structure.ny, structure.nx):
        raise ValueError("labels must have shape (nz, ny, nx)")
    lo, hi = (int(labels.min()), int(labels.max())) if labels.size else (0, -1)
    if lo < 0 or hi >= len(structure.indices):
        raise ValueError("labels must index into `indices`")
    wl = np.asarray(list(wavelengths_um), dtype=np.float64)
    if wl.size == 0:
        raise ValueError("at least one wavelength is required")
    return Handle(structure, tuple(wl.tolist()), _content_hash(structure, labels.astype(np.int32), wl))


def _solver(handle: Handle, wavelength_um: float, precision: str) -> Bidirectional3D:
    if wavelength_um not in handle.wavelengths_um:
        raise ValueError("wavelength was not prepared")
    key = (float(wavelength_um), precision)
    if key not in handle._solvers:
        s = handle.structure
        grid = Grid3D(s.nx, s.ny, s.dx, s.dy, s.dz, wavelength_um)
        mask = absorbing_mask_2d(grid, s.mask_width) if s.mask_width > 0 else None
        handle._solvers[key] = Bidirectional3D(grid, list(s.indices), np.asarray(s.labels), s.index_before, mask,
                                               precision=precision)
    return handle._solvers[key]


# ---- ports ------------------------------------------------------------------------------------

def _port_at(s: Structure, z: float, direction: int, index: complex, wavelength_um: float) -> Port:
    return Port(z, direction, index, wavelength_um, s.nx, s.ny, s.dx, s.dy)


def front_port(handle: Handle, wavelength_um: float, direction: int = 1) -> Port:
    """The plane z0 in the medium before the volume."""
    s = handle.structure
    return _port_at(s, s.z0, direction, s.index_before, wavelength_um)


def back_port(handle: Handle, wavelength_um: float) -> Port:
    """The plane z0 + nz·dz, in the medium of the last slab's centre pixel, direction +1."""
    s = handle.structure
    labels = np.asarray(s.labels)
    exit_label = int(labels[labels.shape[0] - 1, s.ny // 2, s.nx // 2])
    return _port_at(s, s.z0 + labels.shape[0] * s.dz, 1, s.indices[exit_label], wavelength_um)


# ---- solving ----------------------------------------------------------------------------------

def apply(handle: Handle, incident: PortField, iterations: int = 3, backend: str = "auto",
          precision: str = "fp64") -> dict[str, PortField]:
    """Propagate a field incident on the front port; returns {"front": reflected, "back": transmitted}."""
    p, s = incident.port, handle.structure
    on_front = p.direction == 1 and abs(p.z - s.z0) <= 1e-15 and complex(p.index) == complex(s.index_before)
    if not on_front:
        raise ValueError("incident field must be on the front port, direction +1")
    if (p.nx, p.ny, p.dx, p.dy) != (s.nx, s.ny, s.dx, s.dy):
        raise ValueError("port grid differs from the structure grid")
    result = fast.run_3d(_solver(handle, p.wavelength_um, precision), np.asarray(incident.field), iterations,
                         backend=backend)
    wl = p.wavelength_um
    return {"front": PortField(front_port(handle, wl, -1), result.reflected),
            "back": PortField(back_port(handle, wl), result.transmitted)}


def estimate(handle: Handle, iterations: int = 3, backend: str = "auto") -> dict[str, float]:
    """Run time (s) and device memory (bytes) of one `apply`."""
    s = handle.structure
    labels = np.asarray(s.labels)
    nz = labels.shape[0]
    passes = 2 * iterations if iterations > 0 else 1
    points = s.nx * s.ny * nz
    pairs = len({(int(a), int(b)) for a, b in zip(labels[:-1].ravel(), labels[1:].ravel())}) if nz > 1 else 1
    interfaces = int(np.sum(labels[:-1] != labels[1:])) if nz > 1 else 0
    return {
        "seconds": SECONDS_PER_POINT_PASS["gpu"] * points * passes,
        "memory_bytes": float(memory_estimate(s.nx, s.ny, interfaces, pairs, nz=0)),
    }


def port_power(port_field: PortField) -> float:
    """Poynting flux through the port (Parseval sum of Re(k_z)|E|² with E_z from transversality)·dx·dy/N."""
    p = port_field.port
    k0 = 2.0 * np.pi / (p.wavelength_um * 1e-3)
    return _core.power_flux(np.asarray(port_field.field), p.dy, p.dx, k0, p.index, p.direction)


def plane_wave(port: Port, mx: int, my: int, e_perp: tuple[complex, complex]) -> PortField:
    """Plane wave on the port grid with k⊥ = 2π(mx/(nx dx), my/(ny dy)), phase 1 at (0, 0)."""

    def axis(m: int, n: int) -> np.ndarray:
        # e^{i k x_j} with k = 2π m/(n d) and x_j = (j − n//2) d
        return np.exp(2j * np.pi * m * (np.arange(n) - n // 2) / n)

    phase = np.outer(axis(my, port.ny), axis(mx, port.nx))
    return PortField(port, phase[:, :, None] * np.asarray(e_perp, dtype=np.complex128))


__all__ = ["Handle", "Port", "PortField", "Structure", "apply", "back_port", "describe", "estimate", "front_port",
           "plane_wave", "port_power", "prepare"]
