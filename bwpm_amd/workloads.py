"""Synthetic inputs of the benchmark configurations (seeded, reproducible).

Units follow the reference: lengths in mm, wavelengths in µm, k in 1/mm. A Gaussian beam is the
amplitude exp(−(r/w₀)²) on the centred grid x_m = (m − nx//2)·dx (tools/reference_cases.py:66,172).
Labels of tilted interfaces follow the pixel-centre rule (src/bwpm/geometry.py:24-35): slab j is
labelled by the medium at its centre z = (j + ½)·dz.

- `plain_2d`: 2D scalar (TE) BWPM, single planar interface 1.0 → 1.5 at mid-depth (config 1).
- `tilted_3d`: 3D vector BWPM through a planar interface tilted about x (config 2 geometry, any
  transverse size and slab count).
- `born_3d`: homogeneous host with Gaussian-correlated δn (config 3 geometry).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np

from bwpm_amd.specs import Above, FluctuationSynthesis, Geometry, Uniform


@dataclass
class Problem:
    """One solve: grid, labelled volume, input field and run parameters."""

    name: str
    ny: int
    nx: int
    dy: float
    dx: float
    dz: float
    wavelength_um: float
    indices: list[complex]
    labels: np.ndarray  # (nz, ny, nx) int32
    index_before: complex
    field: np.ndarray  # (ny, nx, 2) complex128
    iterations: int
    mask: np.ndarray | None = None
    fluctuation: np.ndarray | None = None
    meta: dict[str, Any] = field(default_factory=dict)

    @property
    def k0(self) -> float:
        return float(2.0 * np.pi / (self.wavelength_um * 1e-3))

    @property
    def nz(self) -> int:
        return int(self.labels.nz) if isinstance(self.labels, (Uniform, Geometry)) else int(self.labels.shape[0])

    def label_volume(self) -> np.ndarray:
        """The (nz, ny, nx) int32 labels, also of a uniform or geometry volume (made on the host;
        a `Geometry` through the library's own label code)."""
        if isinstance(self.labels, Uniform):
            return np.full((self.labels.nz, self.ny, self.nx), self.labels.label, dtype=np.int32)
        if isinstance(self.labels, Geometry):
            from bwpm_amd import _core

            return _core.geometry_labels(self.labels, self.ny, self.nx, self.dy, self.dx, self.dz, len(self.indices))
        return np.asarray(self.labels, dtype=np.int32)

    @property
    def updates(self) -> int:
        """Grid-point·slab updates of the solve: nx·ny·nz per pass, 2 passes per iteration."""
        passes = 2 * self.iterations if self.iterations > 0 else 1
        return self.nx * self.ny * self.nz * passes

    def solver_args(self) -> tuple:
        return (self.ny, self.nx, self.dy, self.dx, self.dz, self.k0, self.indices, self.labels,
                self.index_before, self.mask)


def centred(n: int, d: float) -> np.ndarray:
    return (np.arange(n) - n // 2) * d


def plain_2d(ny: int = 1024, nz: int = 200, iterations: int = 3) -> Problem:
    """Config 1: 1D transverse grid, dy = λ/8, λ = 0.633 µm, TE Gaussian w₀ = 5 µm at normal
    incidence, interface n = 1.0 → 1.5 at mid-depth, dz = λ/4."""
    lam = 0.633
    dy, dz = lam * 1e-3 / 8, lam * 1e-3 / 4
    labels = np.zeros((nz, ny, 1), dtype=np.int32)
    labels[nz // 2:] = 1
    y = centred(ny, dy)
    f = np.zeros((ny, 1, 2), dtype=np.complex128)
    f[:, 0, 0] = np.exp(-(y / 5e-3) ** 2)
    return Problem("plain_2d", ny, 1, dy, 1.0, dz, lam, [1.0 + 0j, 1.5 + 0j], labels, 1.0 + 0j, f, iterations)


def tilted_3d(n: int = 256, nz: int = 512, iterations: int = 4, wavelength_um: float = 1.55,
              tilt_deg: float = 10.0, n1: complex = 1.0, n2: complex = 1.46, polarisation: str = "x",
              waist_mm: float = 8e-3, geometry: bool = False, absorber_wavelengths: float = 0.0) -> Problem:
    """Config 2 geometry: dx = dy = λ/6, dz = λ/4, interface z = z_mid + y·tan(tilt) between n1 and
    n2 (pixel-centre rule), E_x- or E_y-polarised Gaussian of waist w₀ at plane 0. With `geometry`
    the labels are a `Geometry` made inside the library (the same volume, no host array). With
    `absorber_wavelengths` > 0 the transverse edges carry the reference's absorbing zone of that
    many wavelengths (super-Gaussian of order 8, `volume3d.absorbing_mask_2d`); its validity rule
    asks for ≥ 20. Without it the periodic window lets waves refracted at the tilted interface
    circulate: at 512² the reflected field grows about 400× per iteration (the reference's own core
    gives the same fields), while 256² stays bounded."""
    lam = wavelength_um
    dx, dz = lam * 1e-3 / 6, lam * 1e-3 / 4
    y = centred(n, dx)
    x = centred(n, dx)
    tan = np.tan(np.deg2rad(tilt_deg))
    if geometry:
        labels = Geometry(nz, (Above(1, (nz // 2) * dz, 0.0, float(tan)),))
    else:
        zc = (np.arange(nz) + 0.5) * dz
        surface = (nz // 2) * dz + y * tan
        lab2 = (zc[:, None] > surface[None, :]).astype(np.int32)  # (nz, ny)
        labels = np.ascontiguousarray(np.broadcast_to(lab2[:, :, None], (nz, n, n)))
    r2 = x[None, :] ** 2 + y[:, None] ** 2
    f = np.zeros((n, n, 2), dtype=np.complex128)
    f[..., 0 if polarisation == "x" else 1] = np.exp(-r2 / waist_mm ** 2)
    mask = None
    if absorber_wavelengths > 0:
        from bwpm_amd.volume3d import Grid3D, absorbing_mask_2d

        mask = absorbing_mask_2d(Grid3D(n, n, dx, dx, dz, lam), absorber_wavelengths * lam * 1e-3)
    return Problem(f"tilted_3d_{n}x{n}x{nz}", n, n, dx, dx, dz, lam, [complex(n1), complex(n2)], labels,
                   complex(n1), f, iterations, mask=mask,
                   meta={"tilt_deg": tilt_deg, "polarisation": polarisation, "absorber_wavelengths": absorber_wavelengths})


def gaussian_correlated(shape: tuple[int, int, int], spacing_mm: tuple[float, float, float], corr_mm: float,
                        sigma: float, seed: int) -> np.ndarray:
    """Standard-normal noise filtered in k-space by exp(−|q|² l_c²/4) and scaled to std σ
    (the construction of tests/test_fluctuation_3d.py:34-41)."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    dz, dy, dx = spacing_mm
    qz = 2 * np.pi * np.fft.fftfreq(nz, dz)
    qy = 2 * np.pi * np.fft.fftfreq(ny, dy)
    qx = 2 * np.pi * np.fft.fftfreq(nx, dx)
    q2 = qz[:, None, None] ** 2 + qy[None, :, None] ** 2 + qx[None, None, :] ** 2
    noise = rng.standard_normal(shape)
    values = np.real(np.fft.ifftn(np.fft.fftn(noise) * np.exp(-q2 * corr_mm ** 2 / 4)))
    return sigma * values / values.std()


def born_3d(n: int = 512, nz: int = 1024, iterations: int = 6, sigma: float = 0.01, corr_um: float = 1.0,
            seed: int = 3, wavelength_um: float = 0.8, step_um: float = 0.1, angle_deg: float = 5.0) -> Problem:
    """Config 3 geometry: host n = 1.5, Gaussian-correlated δn, plane wave on the grid frequency
    nearest the incidence angle, E_x polarised."""
    lam = wavelength_um
    d = step_um * 1e-3
    host = 1.5
    labels = np.zeros((nz, n, n), dtype=np.int32)
    delta = gaussian_correlated((nz, n, n), (d, d, d), corr_um * 1e-3, sigma, seed).astype(np.complex128)
    k0 = 2 * np.pi / (lam * 1e-3)
    dk = 2 * np.pi / (n * d)
    m = int(round(host * k0 * np.sin(np.deg2rad(angle_deg)) / dk))
    x = centred(n, d)
    f = np.zeros((n, n, 2), dtype=np.complex128)
    f[..., 0] = np.exp(1j * m * dk * x)[None, :]
    return Problem(f"born_3d_{n}x{n}x{nz}", n, n, d, d, d, lam, [complex(host)], labels, complex(host), f,
                   iterations, fluctuation=delta, meta={"m": m, "sigma": sigma})


def born_3d_synth(n: int = 512, nz: int = 1024, iterations: int = 6, sigma: float = 0.01, corr_um: float = 1.0,
                  seed: int = 3, wavelength_um: float = 0.8, step_um: float = 0.1, angle_deg: float = 5.0,
                  host: float = 1.5, dz_um: float | None = None, absorber_wavelengths: float = 0.0) -> Problem:
    """Config 3 geometry with the δn generated on the device (no (nz, ny, nx) arrays on the host):
    a uniform host volume and a FluctuationSynthesis of the same σ and correlation length. The slab
    thickness is `dz_um` (default: the transverse step); `absorber_wavelengths` > 0 adds the
    reference's absorbing zone at the transverse edges."""
    lam = wavelength_um
    d = step_um * 1e-3
    dz = (dz_um if dz_um is not None else step_um) * 1e-3
    k0 = 2 * np.pi / (lam * 1e-3)
    dk = 2 * np.pi / (n * d)
    m = int(round(host * k0 * np.sin(np.deg2rad(angle_deg)) / dk))
    x = centred(n, d)
    f = np.zeros((n, n, 2), dtype=np.complex128)
    f[..., 0] = np.exp(1j * m * dk * x)[None, :]
    mask = None
    if absorber_wavelengths > 0:
        from bwpm_amd.volume3d import Grid3D, absorbing_mask_2d

        mask = absorbing_mask_2d(Grid3D(n, n, d, d, dz, lam), absorber_wavelengths * lam * 1e-3)
    return Problem(f"born_3d_synth_{n}x{n}x{nz}", n, n, d, d, dz, lam, [complex(host)], Uniform(nz, 0),
                   complex(host), f, iterations, mask=mask,
                   fluctuation=FluctuationSynthesis(sigma, corr_um * 1e-3, seed),
                   meta={"m": m, "sigma": sigma, "absorber_wavelengths": absorber_wavelengths})


def uneven_batch(count: int = 64, seed0: int = 1000, sizes: tuple[int, ...] = (128, 256, 384, 512),
                 nz_range: tuple[int, int] = (100, 1000), iterations: int = 1) -> list[Problem]:
    """Config 4: `count` independent Born solves of uneven sizes, one seed per solve (1000 + i):
    n from `sizes`, nz uniform in `nz_range`, λ uniform in 0.4 … 1.6 µm, incidence 0 … 30° on the
    nearest grid frequency, σ uniform in 0.005 … 0.02 with a 1 µm correlation length in a host of
    n = 1.5, dx = dy = λ/6, dz = λ/4; δn generated on the device."""
    out = []
    for i in range(count):
        rng = np.random.default_rng(seed0 + i)
        n = int(rng.choice(sizes))
        nz = int(rng.integers(nz_range[0], nz_range[1] + 1))
        lam = float(rng.uniform(0.4, 1.6))
        theta = float(rng.uniform(0.0, 30.0))
        sigma = float(rng.uniform(0.005, 0.02))
        dx, dz = lam * 1e-3 / 6, lam * 1e-3 / 4
        k0 = 2 * np.pi / (lam * 1e-3)
        dk = 2 * np.pi / (n * dx)
        m = int(round(1.5 * k0 * np.sin(np.deg2rad(theta)) / dk))
        x = centred(n, dx)
        f = np.zeros((n, n, 2), dtype=np.complex128)
        f[..., 0] = np.exp(1j * m * dk * x)[None, :]
        out.append(Problem(f"uneven_{i}_{n}x{n}x{nz}", n, n, dx, dx, dz, lam, [1.5 + 0j], Uniform(nz, 0), 1.5 + 0j,
                           f, iterations, fluctuation=FluctuationSynthesis(sigma, 1e-3, seed0 + i),
                           meta={"m": m, "sigma": sigma, "theta_deg": theta}))
    return out
