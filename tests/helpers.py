"""Input builders shared by the tests (seeded volumes and fields of the reference's tests)."""

from __future__ import annotations

import numpy as np

from oracle import bwpm_oracle as O

SILVER = 0.213 + 6.697j


def k0_of(wavelength_um: float) -> float:
    return float(2.0 * np.pi / (wavelength_um * 1e-3))


def super_gauss_mask(n: int, d: float, width: float, order: int = 8, edge: float = 1e-6) -> np.ndarray:
    """1D super-Gaussian edge absorber on a centred grid (src/bwpm/volume3d.py:70-82 semantics)."""
    a = np.abs((np.arange(n) - n // 2) * d)
    if width == 0:
        return np.ones_like(a)
    u = np.clip((a - (float(np.max(a)) - width)) / width, 0.0, 1.0)
    return edge ** (u ** order)


def mask_2d(ny: int, nx: int, dy: float, dx: float, width: float, order: int = 8) -> np.ndarray:
    return super_gauss_mask(ny, dy, width, order)[:, None] * super_gauss_mask(nx, dx, width, order)[None, :]


def band_limited_field(rng, ny, nx, dy, dx, k0, limit=1.2):
    """Random (E_x, E_y) with |k⊥| < limit·k0 (tests/test_core_solver.py:108-114)."""
    field = rng.normal(size=(ny, nx, 2)) + 1j * rng.normal(size=(ny, nx, 2))
    kx = O.fftfreq(nx, dx)[None, :]
    ky = O.fftfreq(ny, dy)[:, None]
    sp = np.fft.fft2(field, axes=(0, 1))
    sp[np.sqrt(kx ** 2 + ky ** 2) >= limit * k0] = 0.0
    return np.fft.ifft2(sp, axes=(0, 1))


def random_volume(seed: int):
    """Random 3D volume with four materials incl. silver and a mixed slab (test_core_solver.py:99-127)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = (15, 12, 9) if seed % 2 else (8, 16, 11)
    dx, dy, dz, lam = 0.04e-3, 0.05e-3, 0.03e-3, 0.6
    k0 = k0_of(lam)
    labels = rng.integers(0, 4, size=(nz, 1, 1)) * np.ones((1, ny, nx), dtype=np.int64)
    labels[nz // 2] = rng.integers(0, 4, size=(ny, nx))
    indices = [1.0 + 0j, 1.5 + 0j, 2.1 + 0.01j, SILVER]
    field = band_limited_field(rng, ny, nx, dy, dx, k0)
    mask = mask_2d(ny, nx, dy, dx, 2 * dx) if seed == 2 else None
    return dict(ny=ny, nx=nx, dy=dy, dx=dx, dz=dz, k0=k0, indices=indices, labels=labels,
                index_before=1.2 + 0j, mask=mask, field=field)


def fixture_2d_args(d: dict):
    """Map a 2D fixture onto the 3D core with nx = 1, dx = 1 (src/bwpm/fast.py:120-142)."""
    p, i = d["parameters"], d["inputs"]
    labels = np.asarray(i["labels"])[:, :, None]
    mask = i.get("mask")
    fl = i.get("fluctuation")
    return dict(ny=p["ny"], nx=1, dy=p["dy_mm"], dx=1.0, dz=p["dz_mm"], k0=k0_of(p["wavelength_um"]),
                indices=i["indices"], labels=labels, index_before=i["index_before"],
                mask=None if mask is None else np.asarray(mask)[:, None],
                fluctuation=None if fl is None else np.asarray(fl)[:, :, None],
                field=np.asarray(i["field"])[:, None, :])


def fixture_3d_args(d: dict):
    p, i = d["parameters"], d["inputs"]
    return dict(ny=p["ny"], nx=p["nx"], dy=p["dy_mm"], dx=p["dx_mm"], dz=p["dz_mm"],
                k0=k0_of(p["wavelength_um"]), indices=i["indices"], labels=i["labels"],
                index_before=i["index_before"], mask=i.get("mask"), fluctuation=i.get("fluctuation"),
                field=i["field"])


def kz_branch(n: complex, k0: float, kp2: float) -> complex:
    q = complex(np.sqrt(complex(n * n * k0 * k0 - kp2)))
    return -q if (q.imag < 0 or (q.imag == 0 and q.real < 0)) else q


def transfer_rt(media, before, kp2, k0, dz, pol):
    """(r, t) of the tangential field of a planar stack of slabs of thickness dz, from the layer
    transfer matrices (independent of the slab-stepping method)."""
    layers = [before, *media]
    kzs = [kz_branch(n, k0, kp2) for n in layers]
    eta = [k if pol == "s" else n * n * k0 * k0 / k for n, k in zip(layers, kzs)]
    ap, am = 1.0 + 0j, 0.0 + 0j
    for j in range(len(layers) - 1, 0, -1):
        et, ht = ap + am, eta[j] * (ap - am)
        lp, lm = 0.5 * (et + ht / eta[j - 1]), 0.5 * (et - ht / eta[j - 1])
        if j == 1:
            ap, am = lp, lm
        else:
            ap, am = lp * np.exp(-1j * kzs[j - 1] * dz), lm * np.exp(1j * kzs[j - 1] * dz)
    return complex(am / ap), complex(np.exp(1j * kzs[-1] * dz) / ap)
