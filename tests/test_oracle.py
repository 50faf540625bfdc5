"""Pinning of the FP64 oracle (CPU only): the reference's golden fixtures, the reference's own CPU
core compiled from its sources (oracle/_ref, when built), and closed forms from the physics —
an independent vector Fresnel construction, power balance, Brewster, the normal-incidence limit,
the homogeneous plane wave, the transfer matrix of a planar stack, the Airy reflectance of a slab,
exact absorption e^{−2k₀κL} and flux conservation of the Cayley Born step. Planted mistakes in the
oracle must be caught by these checks.
"""

from __future__ import annotations

import importlib.util
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT, load_golden
from helpers import SILVER, band_limited_field, fixture_2d_args, fixture_3d_args, k0_of, mask_2d, random_volume
from oracle import bwpm_oracle as O

K0 = 2 * np.pi / 0.5e-3


def _ref_core():
    so = sorted((ROOT / "oracle" / "_ref").glob("_core*.so")) if (ROOT / "oracle" / "_ref").exists() else []
    if not so:
        pytest.skip("reference core not built (oracle/build_ref.sh)")
    spec = importlib.util.spec_from_file_location("_core", so[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _solver(a: dict) -> O.Solver:
    return O.Solver(a["ny"], a["nx"], a["dy"], a["dx"], a["dz"], a["k0"], a["indices"], a["labels"],
                    a["index_before"], a.get("mask"), fluctuation=a.get("fluctuation"))


# ---- k grid and interface ----------------------------------------------------------------------

@pytest.mark.parametrize(("n", "d"), [(1, 1.0), (12, 0.05e-3), (15, 0.04e-3), (64, 0.5e-3 / 16), (1024, 7.9e-5)])
def test_fftfreq_is_numpy_rounding(n: int, d: float) -> None:
    assert np.array_equal(O.fftfreq(n, d), 2.0 * np.pi * np.fft.fftfreq(n, d=d))


def test_interface_fixture() -> None:
    d = load_golden("interface_matrices")
    kx, ky = d["inputs"]["k_x_per_mm"], d["inputs"]["k_y_per_mm"]
    k0, tol = d["parameters"]["k0_per_mm"], d["tolerance"]
    for j, (n1, n2) in enumerate(d["inputs"]["index_pairs"]):
        s = n1.real
        assert np.allclose(O.transmission_matrix(kx * s, ky * s, n1, n2, k0), d["expected"]["T"][j],
                           rtol=tol["rel"], atol=tol["abs"])
        assert np.allclose(O.reflection_matrix(kx * s, ky * s, n1, n2, k0), d["expected"]["R"][j],
                           rtol=tol["rel"], atol=tol["abs"])


def _fresnel_3d(e_perp: np.ndarray, kx: float, ky: float, n1: complex, n2: complex):
    """Transmitted and reflected (E_x, E_y) from the full 3-vector field: the incident wave is
    completed by k·E = 0, split on the s/p basis of each wave, and scaled by the angle forms of
    the Fresnel coefficients (s along k×z, p = k̂×s)."""
    kp2 = kx * kx + ky * ky

    def kz(n):
        q = np.sqrt(complex(n * n * K0 * K0 - kp2))
        return -q if (q.imag < 0 or (q.imag == 0 and q.real < 0)) else q

    k1, k2 = kz(n1), kz(n2)
    kin, kt, kr = np.array([kx, ky, k1]), np.array([kx, ky, k2]), np.array([kx, ky, -k1])
    e = np.array([e_perp[0], e_perp[1], -(kx * e_perp[0] + ky * e_perp[1]) / k1])
    s = np.array([ky, -kx, 0.0]) / np.sqrt(kp2)
    p_in, p_t, p_r = (np.cross(k / (n * K0), s) for k, n in ((kin, n1), (kt, n2), (kr, n1)))
    c1, c2 = k1 / (n1 * K0), k2 / (n2 * K0)
    rs = (n1 * c1 - n2 * c2) / (n1 * c1 + n2 * c2)
    ts = 2 * n1 * c1 / (n1 * c1 + n2 * c2)
    rp = (n2 * c1 - n1 * c2) / (n2 * c1 + n1 * c2)
    tp = 2 * n1 * c1 / (n2 * c1 + n1 * c2)
    a_s, a_p = e @ s, e @ p_in
    return (ts * a_s * s + tp * a_p * p_t)[:2], (rs * a_s * s + rp * a_p * p_r)[:2]


@pytest.mark.parametrize(("n1", "n2"), [(1.0, 1.5), (1.5, 1.0), (1.0, SILVER), (1.505, SILVER), (1.5, 1.511)])
def test_matrices_equal_vector_fresnel(n1: complex, n2: complex) -> None:
    rng = np.random.default_rng(7)
    theta = rng.uniform(0.0, np.deg2rad(85.0), 200)
    phi = rng.uniform(0.0, 2 * np.pi, 200)
    for th, ph in zip(theta, phi):
        kx, ky = np.real(n1) * K0 * np.sin(th) * np.cos(ph), np.real(n1) * K0 * np.sin(th) * np.sin(ph)
        e = rng.normal(size=2) + 1j * rng.normal(size=2)
        ref_t, ref_r = _fresnel_3d(e, kx, ky, n1, n2)
        t = O.transmission_matrix(np.array(kx), np.array(ky), n1, n2, K0)
        r = O.reflection_matrix(np.array(kx), np.array(ky), n1, n2, K0)
        assert np.max(np.abs(t @ e - ref_t)) < 1e-13 * np.max(np.abs(e))
        assert np.max(np.abs(r @ e - ref_r)) < 1e-13 * np.max(np.abs(e))


def test_normal_incidence_and_power_balance() -> None:
    for n1, n2 in ((1.0, 1.5), (1.5, 1.0), (1.0, SILVER)):
        z = np.zeros(1)
        assert np.allclose(O.transmission_matrix(z, z, n1, n2, K0)[0], 2 * n1 / (n1 + n2) * np.eye(2))
        assert np.allclose(O.reflection_matrix(z, z, n1, n2, K0)[0], (n1 - n2) / (n1 + n2) * np.eye(2))
    # lossless, propagating: reflected + transmitted Poynting flux = incident flux
    k = np.linspace(-0.9, 0.9, 41) * K0 / np.sqrt(2)
    kx, ky = np.meshgrid(k, k)
    t = O.transmission_matrix(kx, ky, 1.0, 1.5, K0)
    r = O.reflection_matrix(kx, ky, 1.0, 1.5, K0)
    e = np.random.default_rng(3).normal(size=(*kx.shape, 2)) + 0j
    kz1 = O.longitudinal_wavenumber(1.0, K0, kx * kx + ky * ky)
    kz2 = O.longitudinal_wavenumber(1.5, K0, kx * kx + ky * ky)

    def flux(f, kz, sign):
        ez = -sign * (kx * f[..., 0] + ky * f[..., 1]) / kz
        return (np.sum(np.abs(f) ** 2, -1) + np.abs(ez) ** 2) * kz.real

    et, er = np.einsum("...ij,...j->...i", t, e), np.einsum("...ij,...j->...i", r, e)
    assert np.max(np.abs((flux(er, kz1, -1) + flux(et, kz2, 1)) / flux(e, kz1, 1) - 1)) < 1e-13


def test_brewster_kills_p_reflection() -> None:
    th = np.arctan(1.5)
    kx = np.array([K0 * np.sin(th)])
    r = O.reflection_matrix(kx, np.zeros(1), 1.0, 1.5, K0)[0]
    assert abs(r[0, 0]) < 1e-15  # p along x for k⊥ ∥ x
    assert abs(r[1, 1]) > 0.1


def test_grazing_rule_zeroes_cancellation_noise() -> None:
    for n in (1.0, 1.5):
        k = n * K0
        for rel in (0.0, 3e-16, -3e-16):
            kz = O.longitudinal_wavenumber(n, K0, np.array([(k * (1 + rel)) ** 2]))
            assert kz[0] == 1j * 1e-12 * K0


# ---- the solver on the reference's fixtures ----------------------------------------------------

def _close(a, b, tol):
    assert a.shape == b.shape
    assert np.allclose(a, b, rtol=tol["rel"], atol=tol["abs"])


def test_fixture_fpwpm2d_interface() -> None:
    d = load_golden("fpwpm2d_interface")
    a = fixture_2d_args(d)
    nz = a["labels"].shape[0]
    out = _solver(a).run(a["field"], 0, list(range(nz + 1)))
    _close(np.stack([out["forward"][p][:, 0] for p in range(nz + 1)]), d["expected"]["planes"], d["tolerance"])


def test_fixture_bwpm2d_tilted_silver_and_residual() -> None:
    d = load_golden("bwpm2d_tilted_silver")
    a = fixture_2d_args(d)
    nz = a["labels"].shape[0]
    out = _solver(a).run(a["field"], 2, list(range(nz + 1)))
    _close(np.stack([out["forward"][p][:, 0] for p in range(nz + 1)]), d["expected"]["forward"], d["tolerance"])
    _close(np.stack([out["backward"][p][:, 0] for p in range(nz + 1)]), d["expected"]["backward"], d["tolerance"])
    p = d["parameters"]
    last = complex(d["inputs"]["indices"][int(d["inputs"]["labels"][-1][p["ny"] // 2])])
    res = O.energy_residual_2d(d["inputs"]["field"], [t[:, 0] for t in out["transmitted_iter"]],
                               [r[:, 0] for r in out["reflected_iter"]], p["ny"], p["dy_mm"], a["k0"], 1.0, last)
    assert np.allclose(res, d["expected"]["energy_residual"], rtol=1e-10, atol=1e-12)


def test_fixture_bwpm2d_slab_exact_and_airy() -> None:
    d = load_golden("bwpm2d_slab_exact")
    a = fixture_2d_args(d)
    out = _solver(a).run(a["field"], 40)
    _close(out["transmitted"][:, 0], d["expected"]["transmitted"], d["tolerance"])
    _close(out["reflected"][:, 0], d["expected"]["reflected"], d["tolerance"])
    # independent: the Airy reflectance of the slab, TE component
    g = d["parameters"]
    k0 = k0_of(g["wavelength_um"])
    ky = 2 * np.pi * 2 / (g["ny"] * g["dy_mm"])
    k1 = np.sqrt(complex(k0 * k0 - ky * ky))
    k2 = np.sqrt(complex(2.25 * k0 * k0 - ky * ky))
    r12 = (k1 - k2) / (k1 + k2)
    ph = np.exp(2j * k2 * 50 * g["dz_mm"])
    r_airy = (r12 - r12 * ph) / (1 - r12 * r12 * ph) * np.exp(2j * k1 * 5 * g["dz_mm"])
    r = complex(np.mean(out["reflected"][:, 0, 0] / d["inputs"]["field"][:, 0]))
    assert abs(r - r_airy) < 1e-12
    assert abs(r_airy - d["reference"]["r_te_airy"]) < 1e-13


def test_fixture_bwpm3d_tilted() -> None:
    d = load_golden("bwpm3d_tilted")
    a = fixture_3d_args(d)
    nz = a["labels"].shape[0]
    keep = (0, 10, nz)
    out = _solver(a).run(a["field"], 2, keep)
    e = d["expected"]
    _close(out["transmitted"], e["transmitted"], d["tolerance"])
    _close(out["reflected"], e["reflected"], d["tolerance"])
    _close(np.stack([out["forward"][q] for q in keep]), e["forward_planes"], d["tolerance"])
    _close(np.stack([out["backward"][q] for q in keep]), e["backward_planes"], d["tolerance"])


def test_fixture_fpwpm2d_fluctuation() -> None:
    d = load_golden("fpwpm2d_fluctuation")
    a = fixture_2d_args(d)
    nz = a["labels"].shape[0]
    out = _solver(a).run(a["field"], 0, list(range(nz + 1)))
    _close(np.stack([out["forward"][p][:, 0] for p in range(nz + 1)]), d["expected"]["planes"], d["tolerance"])


def test_fixture_bwpm3d_fluctuation() -> None:
    d = load_golden("bwpm3d_fluctuation")
    a = fixture_3d_args(d)
    nz = a["labels"].shape[0]
    keep = (0, 8, nz)
    out = _solver(a).run(a["field"], 1, keep)
    e = d["expected"]
    _close(out["transmitted"], e["transmitted"], d["tolerance"])
    _close(out["reflected"], e["reflected"], d["tolerance"])
    _close(np.stack([out["forward"][q] for q in keep]), e["forward_planes"], d["tolerance"])
    _close(np.stack([out["backward"][q] for q in keep]), e["backward_planes"], d["tolerance"])


# ---- closed forms --------------------------------------------------------------------------------

STEP = 1e-3 / 16


def _plane_wave(n: int, mx: int, my: int, e0):
    x = (np.arange(n) - n // 2) * STEP
    kx, ky = 2 * np.pi * mx / (n * STEP), 2 * np.pi * my / (n * STEP)
    ph = np.exp(1j * (kx * x[None, :] + ky * x[:, None]))
    return np.stack([e0[0] * ph, e0[1] * ph], -1), kx, ky


@pytest.mark.parametrize(("mx", "my"), [(0, 0), (3, 1), (-2, 4)])
def test_homogeneous_plane_wave(mx: int, my: int) -> None:
    n, nz, idx = 32, 20, 1.4 + 0.01j
    k0 = k0_of(1.0)
    f, kx, ky = _plane_wave(n, mx, my, (0.3 + 0.1j, -0.7))
    out = O.Solver(n, n, STEP, STEP, STEP, k0, [idx], np.zeros((nz, n, n), int), idx).run(f, 0)
    kz = np.sqrt(complex(idx * idx * k0 * k0 - kx * kx - ky * ky))
    assert np.max(np.abs(out["transmitted"] - f * np.exp(1j * kz * nz * STEP))) < 1e-12


def _transfer(media, before, kp2, k0, dz, pol):
    """(r, t) of the tangential field of a planar stack from the layer transfer matrices."""
    layers = [before, *media]

    def kz(n):
        q = np.sqrt(complex(n * n * k0 * k0 - kp2))
        return -q if (q.imag < 0 or (q.imag == 0 and q.real < 0)) else q

    kzs = [kz(n) for n in layers]
    eta = [k if pol == "s" else n * n * k0 * k0 / k for n, k in zip(layers, kzs)]
    ap, am = 1.0 + 0j, 0.0 + 0j
    for j in range(len(layers) - 1, 0, -1):
        et, ht = ap + am, eta[j] * (ap - am)
        lp, lm = 0.5 * (et + ht / eta[j - 1]), 0.5 * (et - ht / eta[j - 1])
        if j == 1:
            ap, am = lp, lm
        else:
            ap, am = lp * np.exp(-1j * kzs[j - 1] * dz), lm * np.exp(1j * kzs[j - 1] * dz)
    return am / ap, np.exp(1j * kzs[-1] * dz) / ap


@pytest.mark.parametrize(("mx", "my"), [(1, 0), (1, 1), (0, 0)])
def test_planar_stack_against_transfer_matrix(mx: int, my: int) -> None:
    n, k0 = 32, k0_of(1.0)
    media = [1.0 + 0j] * 3 + [1.5 + 0j] * 10 + [2.1 + 0j] * 5 + [SILVER] * 3 + [1.5 + 0j] * 4
    uniq = sorted(set(media), key=lambda c: (c.real, c.imag))
    labels = np.array([np.full((n, n), uniq.index(m)) for m in media])
    e0 = (0.6 + 0.2j, -0.3 + 0.5j)
    f, kx, ky = _plane_wave(n, mx, my, e0)
    out = O.Solver(n, n, STEP, STEP, STEP, k0, uniq, labels, 1.0).run(f, 30)
    kp2 = kx * kx + ky * ky
    if kp2 > 0:
        sd, pd = np.array([-ky, kx]) / np.sqrt(kp2), np.array([kx, ky]) / np.sqrt(kp2)
    else:
        sd, pd = np.array([0.0, 1.0]), np.array([1.0, 0.0])
    rs, ts = _transfer(media, 1.0, kp2, k0, STEP, "s")
    rp, tp = _transfer(media, 1.0, kp2, k0, STEP, "p")
    a_s, a_p = np.array(e0) @ sd, np.array(e0) @ pd
    c = (n // 2, n // 2)
    assert np.max(np.abs(out["transmitted"][c] - (ts * a_s * sd + tp * a_p * pd))) < 1e-12
    assert np.max(np.abs(out["reflected"][c] - (rs * a_s * sd + rp * a_p * pd))) < 1e-12


def _flux(field, n0, k0, d):
    n = field.shape[0]
    kx = O.fftfreq(n, d)[None, :]
    ky = O.fftfreq(n, d)[:, None]
    k2 = (n0 * k0) ** 2
    prop = kx ** 2 + ky ** 2 < k2 * (1 - 1e-12)
    kz = np.sqrt(np.where(prop, k2 - kx ** 2 - ky ** 2, 1.0))
    sp = np.fft.fft2(field, axes=(0, 1))
    ez = -(kx * sp[..., 0] + ky * sp[..., 1]) / kz
    return float(np.sum(np.where(prop, kz * (np.sum(np.abs(sp) ** 2, -1) + np.abs(ez) ** 2), 0.0)))


def test_uniform_absorbing_fluctuation_attenuates_exactly() -> None:
    n, nz, h, kappa = 8, 100, 0.05e-3, 1e-3
    k0 = k0_of(0.8)
    f = np.zeros((n, n, 2), complex)
    f[..., 0] = 1.0
    s = O.Solver(n, n, h, h, h, k0, [1.38], np.zeros((nz, n, n), int), 1.38,
                 fluctuation=np.full((nz, n, n), 1j * kappa))
    out = s.run(f, 0)["transmitted"]
    expected = np.exp(-2 * k0 * kappa * nz * h)
    assert abs(_flux(out, 1.38, k0, h) / _flux(f, 1.38, k0, h) / expected - 1) < 1e-6


def test_real_fluctuation_conserves_flux() -> None:
    n, nz, h = 16, 12, 0.1e-3
    k0 = k0_of(0.8)
    rng = np.random.default_rng(1)
    delta = 0.05 * rng.standard_normal((nz, n, n))
    f = np.zeros((n, n, 2), complex)
    f[..., 0] = 1.0
    out = O.Solver(n, n, h, h, h, k0, [1.38], np.zeros((nz, n, n), int), 1.38, fluctuation=delta).run(f, 0)
    assert abs(_flux(out["transmitted"], 1.38, k0, h) / _flux(f, 1.38, k0, h) - 1) < 1e-8


def test_zero_fluctuation_is_bit_identical() -> None:
    a = fixture_2d_args(load_golden("bwpm2d_tilted_silver"))
    plain = _solver(a).run(a["field"], 2)
    b = dict(a, fluctuation=np.zeros(a["labels"].shape))
    zero = _solver(b).run(b["field"], 2)
    assert np.array_equal(plain["transmitted"], zero["transmitted"])
    assert np.array_equal(plain["reflected"], zero["reflected"])


def test_gmres_solves_a_known_system() -> None:
    rng = np.random.default_rng(4)
    a = np.eye(40) + 0.1 * (rng.normal(size=(40, 40)) + 1j * rng.normal(size=(40, 40)))
    b = rng.normal(size=40) + 1j * rng.normal(size=40)
    ok, x, its = O.gmres(lambda v: a @ v, b, np.zeros(40, complex), 1e-12, 8, 100)
    assert ok and its > 8  # restarted
    assert np.allclose(x, np.linalg.solve(a, b), rtol=1e-10, atol=1e-12)
    ok, _, _ = O.gmres(lambda v: a @ v, b, np.zeros(40, complex), 1e-12, 2, 1)
    assert not ok


def test_host_rule_rejects_two_hosts() -> None:
    labels = np.zeros((2, 4, 4), int)
    labels[0, :, :2] = 1
    fl = np.zeros((2, 4, 4))
    fl[0] = 0.01
    with pytest.raises(ValueError, match="host"):
        O.Solver(4, 4, 1e-4, 1e-4, 1e-4, 1e4, [1.0, 1.5], labels, 1.0, fluctuation=fl)


# ---- the reference's compiled CPU core ---------------------------------------------------------

@pytest.mark.parametrize("seed", [1, 2, 3])
def test_equals_reference_core_random_volumes(seed: int) -> None:
    R = _ref_core()
    a = random_volume(seed)
    nz = a["labels"].shape[0]
    keep = [0, 5, nz]
    mine = _solver(a).run(a["field"], 3, keep)
    ref = R.Solver(a["ny"], a["nx"], a["dy"], a["dx"], a["dz"], a["k0"], np.asarray(a["indices"], complex),
                   a["labels"].astype(np.int32), complex(a["index_before"]), a["mask"], 2).run(a["field"], 3, keep)
    scale = 1e-12 * float(np.max(np.abs(a["field"])))
    for key in ("transmitted", "reflected"):
        assert np.max(np.abs(mine[key] - ref[key])) < scale
    for q in keep:
        assert np.max(np.abs(mine["forward"][q] - ref["forward"][q])) < scale
        assert np.max(np.abs(mine["backward"][q] - ref["backward"][q])) < scale
    assert mine["stored_values"] == ref["stored_values"]


@pytest.mark.parametrize("seed", [3, 4])
def test_equals_reference_core_back_scattering(seed: int) -> None:
    """δn with R1–R3 over three iterations (the path only the core-vs-Python test pinned)."""
    R = _ref_core()
    rng = np.random.default_rng(seed)
    nx, ny, nz = (12, 10, 14) if seed % 2 else (8, 16, 11)
    dx, dz, k0 = 0.05e-3, 0.04e-3, k0_of(0.8)
    labels = np.zeros((nz, ny, nx), dtype=np.int32)
    labels[nz - 3:] = 1
    fl = np.zeros((nz, ny, nx), complex)
    fl[1:nz - 4] = 0.02 * rng.normal(size=(nz - 5, ny, nx)) + 1e-3j
    field = rng.normal(size=(ny, nx, 2)) + 1j * rng.normal(size=(ny, nx, 2))
    mask = mask_2d(ny, nx, dx, dx, 2 * dx) if seed == 4 else None
    args = (ny, nx, dx, dx, dz, k0, np.array([1.38, 1.5 + 0.002j]), labels, 1.38 + 0j, mask)
    mine = O.Solver(*args, fluctuation=fl).run(field, 3, [0, 5, nz])
    ref = R.Solver(*args, 2, fluctuation=fl).run(field, 3, [0, 5, nz])
    scale = 1e-12 * float(np.max(np.abs(field)))
    for key in ("transmitted", "reflected"):
        assert np.max(np.abs(mine[key] - ref[key])) < scale


# ---- planted mistakes must be caught -----------------------------------------------------------

def _fixture_ok() -> bool:
    d = load_golden("bwpm2d_tilted_silver")
    a = fixture_2d_args(d)
    nz = a["labels"].shape[0]
    out = _solver(a).run(a["field"], 2, list(range(nz + 1)))
    got = np.stack([out["backward"][p][:, 0] for p in range(nz + 1)])
    return bool(np.allclose(got, d["expected"]["backward"], rtol=d["tolerance"]["rel"], atol=d["tolerance"]["abs"]))


@pytest.mark.parametrize("mistake", ["reflection_transposed_sign", "bwd_phase_from_s", "no_mask"])
def test_planted_mistakes_fail_the_fixture(mistake: str, monkeypatch: pytest.MonkeyPatch) -> None:
    assert _fixture_ok()
    if mistake == "reflection_transposed_sign":
        orig = O.reflection_matrix
        monkeypatch.setattr(O, "reflection_matrix", lambda *a: -orig(*a))
    elif mistake == "bwd_phase_from_s":
        orig_op = O.Solver.op

        def op(self, kind, r, s):
            return orig_op(self, O.FWD_T if kind == O.BWD_T else kind, r, s)

        monkeypatch.setattr(O.Solver, "op", op)
    else:
        orig_init = O.Solver.__init__

        def init(self, *a, **k):
            orig_init(self, *a, **k)
            self.mask = None

        monkeypatch.setattr(O.Solver, "__init__", init)
    assert not _fixture_ok()


# ---- 3D port flux (oracle.port_power) ----------------------------------------------------------

def _port_plane_wave(nx, ny, dx, dy, mx, my, e_perp):
    x = np.arange(nx) * dx
    y = np.arange(ny) * dy
    kx = 2 * np.pi * mx / (nx * dx)
    ky = 2 * np.pi * my / (ny * dy)
    ph = np.exp(1j * (kx * x[None, :] + ky * y[:, None]))
    return np.stack([e_perp[0] * ph, e_perp[1] * ph], axis=-1), kx, ky


@pytest.mark.parametrize("direction", [1, -1])
def test_port_power_reduces_to_the_2d_flux(direction: int) -> None:
    # a field that does not vary along x (nx = 1): the 3D flux is dx times the 2D flux per unit x
    rng = np.random.default_rng(11)
    ny, dy, dx, k0, n = 64, 0.5e-3 / 8, 0.37e-3, 2 * np.pi / 0.5e-3, 1.3 + 0j
    f = band_limited_field(rng, ny, 1, dy, dx, k0)
    got = O.port_power(f, 1, ny, dx, dy, k0, n, direction)
    want = dx * O.power_flux(f[:, 0, :], ny, dy, k0, n, direction)
    assert abs(got - want) <= 1e-13 * abs(want)


@pytest.mark.parametrize(("mx", "my", "n"), [(0, 0, 1.0), (3, 0, 1.0), (0, 5, 1.46), (4, -3, 1.46), (-6, 2, 2.1)])
def test_port_power_of_plane_waves_is_n_k0_cos_theta(mx: int, my: int, n: float) -> None:
    # a plane wave of unit |E| (E ⊥ k) carries n·k0·cosθ·|E|²·area for s and p polarisation alike;
    # the p wave needs the E_z term with the right wave-vector components
    nx, ny, dx, dy = 24, 20, 0.5e-3 / 6, 0.5e-3 / 5
    k0 = 2 * np.pi / 0.5e-3
    kxv = 2 * np.pi * mx / (nx * dx)
    kyv = 2 * np.pi * my / (ny * dy)
    kp = np.hypot(kxv, kyv)
    kz = np.sqrt((n * k0) ** 2 - kp ** 2)
    cos_t = kz / (n * k0)
    area = nx * dx * ny * dy
    if kp == 0:
        s_dir, p_dir = np.array([0.0, 1.0]), np.array([1.0, 0.0])
    else:
        s_dir, p_dir = np.array([-kyv, kxv]) / kp, np.array([kxv, kyv]) / kp
    a, b = 0.6 - 0.8j, 0.3 + 1.4j
    # s wave: E = a·ŝ (transverse); p wave: |E| = |b|, transverse part b·cosθ along k⊥
    for e_perp, e_total2 in ((s_dir * a, abs(a) ** 2), (p_dir * cos_t * b, abs(b) ** 2)):
        field, _, _ = _port_plane_wave(nx, ny, dx, dy, mx, my, e_perp)
        for direction in (1, -1):
            got = O.port_power(field, nx, ny, dx, dy, k0, n + 0j, direction)
            want = n * k0 * cos_t * e_total2 * area
            assert abs(got - want) <= 1e-12 * want


def test_port_power_of_an_evanescent_wave_is_zero() -> None:
    nx, ny, dx, dy = 16, 16, 0.5e-3 / 4, 0.5e-3 / 4
    k0 = 2 * np.pi / 0.5e-3
    field, kx, ky = _port_plane_wave(nx, ny, dx, dy, 6, 5, (1.0, 0.5j))
    assert np.hypot(kx, ky) > k0
    # only FFT rounding in the propagating modes is left (relative to k0·|E|²·area)
    scale = k0 * 1.25 * nx * dx * ny * dy
    assert abs(O.port_power(field, nx, ny, dx, dy, k0, 1.0 + 0j)) <= 1e-26 * scale


def test_planted_mistake_in_port_power_fails_the_closed_form(monkeypatch: pytest.MonkeyPatch) -> None:
    # E_z from the wrong wave-vector component (k_y·E_x + k_x·E_y) must fail the p-wave check
    orig = O.port_power

    def swapped(field, nx, ny, dx, dy, k0, index, direction=1):
        f = np.asarray(field)
        return orig(f[..., ::-1], nx, ny, dx, dy, k0, index, direction)

    monkeypatch.setattr(O, "port_power", swapped)
    with pytest.raises(AssertionError):
        test_port_power_of_plane_waves_is_n_k0_cos_theta(4, 0, 1.46)
