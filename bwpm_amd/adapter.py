"""Plugin adapter of the GPU engine with the reference's `bwpm.adapter` API (src/bwpm/adapter.py:38-310).

- `describe()`: method, conventions, ports, backends and validity limits as data.
- `prepare(structure, wavelengths)`: checks the structure, computes a content hash.
- `apply(handle, port_field, ...)`: field incident on the front port (+z) → reflected field at the
  front port (−z) and transmitted field at the back port (+z), solved on the GPU.
- `estimate(handle, iterations)`: run time from the measured MI355X update rate and device memory
  from the engine's memory model.
- `port_power(port_field)`: Poynting flux through a port (device FP64 reduction).
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from bwpm_amd import __version__, _core, fast
from bwpm_amd.volume3d import Bidirectional3D, Grid3D, absorbing_mask_2d, memory_estimate

# seconds per grid point and plane pass on one MI355X (single solve, 512² plain slab steps, measured
# by bench.py); the batched lanes of `bwpm_amd.serve` reach several times this rate
SECONDS_PER_POINT_PASS = {"gpu": 1.6e-10}


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
        if np.asarray(self.field).shape != (self.port.ny, self.port.nx, 2):
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
    _solvers: dict[float, Bidirectional3D] = field(default_factory=dict, repr=False)


def _backends() -> list[str]:
    return ["gpu"] if fast.gpu_available() else []


def describe() -> dict[str, Any]:
    """Self-description for a planner."""
    return {
        "name": "BWPM",
        "version": __version__,
        "method": "fast bidirectional vector wave propagation (wende2023), FPWPM steps with Fresnel "
                  "interface matrices, forward-backward iterations; MI355X engine",
        "conventions": {
            "time": "exp(-i omega t), plane waves exp(i(k.r - omega t))",
            "fresnel": "Convention A (Byrnes arXiv:1603.02720v5 eq. 6)",
            "polarisation": "s = TE, p = TM; port fields are transverse (E_x, E_y)",
            "units": {"length": "mm", "wavelength": "um", "k": "1/mm"},
            "grid": "centred, x_m = (m - nx//2) dx; spectra referred to (0, 0)",
            "kz": "Im k_z >= 0; grazing rule |n^2 k0^2 - kperp^2| <= 1e-12 |n^2 k0^2|",
        },
        "ports": [
            {"name": "front", "plane": "z0", "in": "+z", "out": "-z (reflected)"},
            {"name": "back", "plane": "z0 + nz dz", "out": "+z (transmitted)"},
        ],
        "representation": "gridded transverse field",
        "linear": True,
        "reciprocal": True,
        "operator": False,
        "deterministic": {"gpu": "bitwise (fixed reduction order, batch-invariant)"},
        "precision": {"fp64": "complex128 bulk", "fp32": "complex64 bulk, FP64 tables, compensated"},
        "backends": _backends(),
        "validity": [
            "reflections quantitative only at z-normal interfaces; tilted surfaces: direction and phase only",
            "waves trapped by total internal reflection do not converge; limit iterations",
            "absorbing boundary zone >= 20 wavelengths, order 8",
        ],
    }


def prepare(structure: Structure, wavelengths_um: list[float] | tuple[float, ...]) -> Handle:
    """Check the structure and return a handle with a content hash (nothing is computed)."""
    labels = np.asarray(structure.labels)
    if labels.ndim != 3 or labels.shape[1:] != (structure.ny, structure.nx):
        raise ValueError("labels must have shape (nz, ny, nx)")
    if labels.min() < 0 or labels.max() >= len(structure.indices):
        raise ValueError("labels must index into `indices`")
    if not wavelengths_um:
        raise ValueError("at least one wavelength is required")
    h = hashlib.sha256()
    h.update(f"bwpm_amd {__version__}".encode())
    for value in (structure.nx, structure.ny, structure.dx, structure.dy, structure.dz, structure.index_before,
                  structure.z0, structure.mask_width):
        h.update(repr(value).encode())
    h.update(np.asarray(structure.indices, dtype=np.complex128).tobytes())
    h.update(np.ascontiguousarray(labels, dtype=np.int32).tobytes())
    h.update(np.asarray(wavelengths_um, dtype=np.float64).tobytes())
    return Handle(structure, tuple(float(w) for w in wavelengths_um), h.hexdigest())


def _solver(handle: Handle, wavelength_um: float, precision: str) -> Bidirectional3D:
    if wavelength_um not in handle.wavelengths_um:
        raise ValueError("wavelength was not prepared")
    s = handle.structure
    grid = Grid3D(s.nx, s.ny, s.dx, s.dy, s.dz, wavelength_um)
    mask = absorbing_mask_2d(grid, s.mask_width) if s.mask_width > 0 else None
    return Bidirectional3D(grid, list(s.indices), np.asarray(s.labels), s.index_before, mask, precision=precision)


def front_port(handle: Handle, wavelength_um: float, direction: int = 1) -> Port:
    s = handle.structure
    return Port(s.z0, direction, s.index_before, wavelength_um, s.nx, s.ny, s.dx, s.dy)


def back_port(handle: Handle, wavelength_um: float) -> Port:
    s = handle.structure
    labels = np.asarray(s.labels)
    last = s.indices[int(labels[-1, s.ny // 2, s.nx // 2])]
    return Port(s.z0 + labels.shape[0] * s.dz, 1, last, wavelength_um, s.nx, s.ny, s.dx, s.dy)


def apply(handle: Handle, incident: PortField, iterations: int = 3, backend: str = "auto",
          precision: str = "fp64") -> dict[str, PortField]:
    """Propagate a field incident on the front port; returns {"front": reflected, "back": transmitted}."""
    p = incident.port
    s = handle.structure
    if p.direction != 1 or abs(p.z - s.z0) > 1e-15 or complex(p.index) != complex(s.index_before):
        raise ValueError("incident field must be on the front port, direction +1")
    if (p.nx, p.ny, p.dx, p.dy) != (s.nx, s.ny, s.dx, s.dy):
        raise ValueError("port grid differs from the structure grid")
    solver = _solver(handle, p.wavelength_um, precision)
    result = fast.run_3d(solver, np.asarray(incident.field), iterations, backend=backend)
    return {
        "front": PortField(front_port(handle, p.wavelength_um, -1), result.reflected),
        "back": PortField(back_port(handle, p.wavelength_um), result.transmitted),
    }


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
    x = (np.arange(port.nx) - port.nx // 2) * port.dx
    y = (np.arange(port.ny) - port.ny // 2) * port.dy
    kx = 2 * np.pi * mx / (port.nx * port.dx)
    ky = 2 * np.pi * my / (port.ny * port.dy)
    phase = np.exp(1j * (kx * x[None, :] + ky * y[:, None]))
    return PortField(port, np.stack([e_perp[0] * phase, e_perp[1] * phase], axis=-1))


__all__ = ["Handle", "Port", "PortField", "Structure", "apply", "back_port", "describe", "estimate", "front_port",
           "plane_wave", "port_power", "prepare"]
