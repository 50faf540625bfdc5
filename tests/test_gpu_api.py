"""The drop-in API on the GPU: `bwpm_amd.fast`, `fpwpm`, `bidirectional`, `volume3d`, `interface`
and `adapter` behave like the reference's modules — fixtures through the facade, and the
reference's conformance properties (tests/test_adapter_conformance.py): planar-stack transfer
matrix, Airy, energy balance, reciprocity, bitwise determinism, shift equivariance."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden
from helpers import kz_branch, transfer_rt
from oracle import bwpm_oracle as O

pytestmark = pytest.mark.gpu

LAMBDA_UM = 1.0
STEP = LAMBDA_UM * 1e-3 / 16
K0 = 2 * np.pi / (LAMBDA_UM * 1e-3)
SILVER = 0.213 + 6.697j
STACK = [1.0 + 0j] * 2 + [1.5 + 0j] * 9 + [2.1 + 0.02j] * 4 + [1.45 + 0j] * 6 + [1.0 + 0j] * 3


def _close64(a, b):
    assert np.allclose(a, b, rtol=1e-10, atol=1e-11 * float(np.max(np.abs(b)) or 1.0))


def _structure(media, n=32, before=1.0 + 0j):
    from bwpm_amd import adapter

    unique = sorted(set(media), key=lambda c: (c.real, c.imag))
    labels = np.array([np.full((n, n), unique.index(m)) for m in media], dtype=np.int64)
    return adapter.Structure(n, n, STEP, STEP, STEP, tuple(unique), labels, before)


def _prop2d(d):
    from bwpm_amd.fpwpm import Grid2D, Propagator2D

    p, i = d["parameters"], d["inputs"]
    g = Grid2D(p["ny"], p["dy_mm"], p["dz_mm"], p["wavelength_um"])
    return Propagator2D(g, list(i["indices"]), i["labels"], i["index_before"], i.get("mask"),
                        fluctuation=i.get("fluctuation"))


def test_fast_run_2d_fixtures_and_residual() -> None:
    from bwpm_amd import fast

    d = load_golden("fpwpm2d_interface")
    res = fast.run_2d(_prop2d(d), d["inputs"]["field"], iterations=0, keep=True)
    _close64(res.forward, d["expected"]["planes"])
    d = load_golden("bwpm2d_tilted_silver")
    res = fast.run_2d(_prop2d(d), d["inputs"]["field"], iterations=2, keep=True)
    _close64(res.forward, d["expected"]["forward"])
    _close64(res.backward, d["expected"]["backward"])
    assert np.allclose(res.energy_residual, d["expected"]["energy_residual"], rtol=1e-10, atol=1e-12)


def test_propagator_and_bidirectional_classes() -> None:
    from bwpm_amd.bidirectional import Bidirectional2D

    d = load_golden("fpwpm2d_interface")
    prop = _prop2d(d)
    planes = d["expected"]["planes"]
    _close64(prop.run(d["inputs"]["field"], keep=True), planes)
    labels = np.asarray(d["inputs"]["labels"])
    # single steps, including the step that crosses the interface at slab 8
    _close64(prop.step(d["inputs"]["field"], np.full(labels.shape[1], -1), labels[0]), planes[1])
    _close64(prop.step(planes[8], labels[7], labels[8]), planes[9])
    d = load_golden("bwpm2d_slab_exact")
    res = Bidirectional2D(_prop2d(d)).run(d["inputs"]["field"], 40)
    _close64(res.transmitted, d["expected"]["transmitted"])
    with pytest.raises(ValueError):
        Bidirectional2D(_prop2d(d)).run(d["inputs"]["field"], 0)


def test_propagator_step_with_fluctuation_equals_fixture() -> None:
    d = load_golden("fpwpm2d_fluctuation")
    prop = _prop2d(d)
    planes = d["expected"]["planes"]
    labels = np.asarray(d["inputs"]["labels"])
    fl = np.asarray(d["inputs"]["fluctuation"])
    _close64(prop.step(planes[5], labels[4], labels[5], fl[5], 1.38), planes[6])


def test_volume3d_and_interface_modules() -> None:
    from bwpm_amd import interface
    from bwpm_amd.volume3d import Bidirectional3D, Grid3D

    d = load_golden("bwpm3d_tilted")
    p, i = d["parameters"], d["inputs"]
    g = Grid3D(p["nx"], p["ny"], p["dx_mm"], p["dy_mm"], p["dz_mm"], p["wavelength_um"])
    nz = i["labels"].shape[0]
    res = Bidirectional3D(g, list(i["indices"]), i["labels"], i["index_before"], i["mask"]).run(
        i["field"], 2, keep=(0, 10, nz))
    _close64(res.transmitted, d["expected"]["transmitted"])
    _close64(np.stack([res.planes[q][1] for q in (0, 10, nz)]), d["expected"]["backward_planes"])
    kx = np.linspace(-1, 1, 7) * K0
    c = interface.fresnel_coefficients(kx, 0.3 * kx, 1.0, 1.5, K0)
    r_s, r_p, t_s, t_p, kz1, kz2 = O.fresnel_coefficients(kx, 0.3 * kx, 1.0, 1.5, K0)
    for got, want in ((c.r_s, r_s), (c.r_p, r_p), (c.t_s, t_s), (c.t_p, t_p), (c.kz_in, kz1), (c.kz_out, kz2)):
        assert np.allclose(got, want, rtol=1e-13, atol=1e-13 * K0)


def test_power_flux_matches_oracle() -> None:
    from bwpm_amd import fpwpm

    d = load_golden("bwpm2d_tilted_silver")
    prop = _prop2d(d)
    f = d["inputs"]["field"]
    for direction in (1, -1):
        got = fpwpm.power_flux(f, prop.grid, 1.0, direction)
        want = O.power_flux(f, prop.grid.ny, prop.grid.dy, prop.grid.k0, 1.0, direction)
        assert abs(got - want) < 1e-12 * abs(want)


@pytest.mark.parametrize(("mx", "my"), [(0, 0), (1, 0), (1, 1), (-1, 1)])
def test_planar_stack_against_transfer_matrix(mx: int, my: int) -> None:
    from bwpm_amd import adapter

    handle = adapter.prepare(_structure(STACK), [LAMBDA_UM])
    e0 = (0.6 + 0.2j, -0.3 + 0.5j)
    incident = adapter.plane_wave(adapter.front_port(handle, LAMBDA_UM), mx, my, e0)
    out = adapter.apply(handle, incident, iterations=30)
    kx, ky = 2 * np.pi * mx / (32 * STEP), 2 * np.pi * my / (32 * STEP)
    kp2 = kx * kx + ky * ky
    s_dir = np.array([-ky, kx]) / np.sqrt(kp2) if kp2 else np.array([0.0, 1.0])
    p_dir = np.array([kx, ky]) / np.sqrt(kp2) if kp2 else np.array([1.0, 0.0])
    e = np.array(e0)
    r_s, t_s = transfer_rt(STACK, 1.0, kp2, K0, STEP, "s")
    r_p, t_p = transfer_rt(STACK, 1.0, kp2, K0, STEP, "p")
    c = (16, 16)
    assert np.max(np.abs(out["front"].field[c] - (r_s * (e @ s_dir) * s_dir + r_p * (e @ p_dir) * p_dir))) < 1e-12
    assert np.max(np.abs(out["back"].field[c] - (t_s * (e @ s_dir) * s_dir + t_p * (e @ p_dir) * p_dir))) < 1e-12


def test_slab_against_airy() -> None:
    from bwpm_amd import adapter

    media = [1.0 + 0j] * 4 + [1.5 + 0j] * 40 + [1.0 + 0j] * 4
    handle = adapter.prepare(_structure(media), [LAMBDA_UM])
    incident = adapter.plane_wave(adapter.front_port(handle, LAMBDA_UM), 0, 1, (1.0, 0.0))
    out = adapter.apply(handle, incident, iterations=40)
    ky = 2 * np.pi / (32 * STEP)
    k1, k2 = kz_branch(1.0, K0, ky * ky), kz_branch(1.5, K0, ky * ky)
    r12 = (k1 - k2) / (k1 + k2)
    ph = np.exp(2j * k2 * 40 * STEP)
    r_airy = (r12 - r12 * ph) / (1 - r12 * r12 * ph) * np.exp(2j * k1 * 4 * STEP)
    assert abs(out["front"].field[16, 16, 0] - r_airy) < 1e-12


def test_energy_conservation_lossless() -> None:
    from bwpm_amd import adapter

    media = [1.0 + 0j] * 2 + [1.5 + 0j] * 9 + [2.1 + 0j] * 4 + [1.45 + 0j] * 6 + [1.0 + 0j] * 3
    handle = adapter.prepare(_structure(media), [LAMBDA_UM])
    port = adapter.front_port(handle, LAMBDA_UM)
    rng = np.random.default_rng(4)
    field = rng.normal(size=(32, 32, 2)) + 1j * rng.normal(size=(32, 32, 2))
    k = 2 * np.pi * np.fft.fftfreq(32, STEP)
    kxx, kyy = np.meshgrid(k, k)
    sp = np.fft.fft2(field, axes=(0, 1))
    sp[np.sqrt(kxx ** 2 + kyy ** 2) >= 0.9 * K0] = 0.0
    incident = adapter.PortField(port, np.fft.ifft2(sp, axes=(0, 1)))
    out = adapter.apply(handle, incident, iterations=60)
    balance = (adapter.port_power(out["front"]) + adapter.port_power(out["back"])) / adapter.port_power(incident)
    assert abs(balance - 1.0) < 1e-12


def test_reciprocity_of_transmission() -> None:
    from bwpm_amd import adapter

    media = [1.0 + 0j] * 2 + [1.5 + 0j] * 5 + [SILVER] * 1 + [2.1 + 0.05j] * 7 + [1.0 + 0j] * 3
    res = []
    for m in (media, media[::-1]):
        h = adapter.prepare(_structure(m), [LAMBDA_UM])
        res.append(adapter.apply(h, adapter.plane_wave(adapter.front_port(h, LAMBDA_UM), 1, 1, (0.4, 0.7j)), 40))
    assert np.max(np.abs(res[0]["back"].field[16, 16] - res[1]["back"].field[16, 16])) < 1e-12
    assert np.max(np.abs(res[0]["front"].field[16, 16] - res[1]["front"].field[16, 16])) > 1e-3


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_determinism_is_bitwise(precision: str) -> None:
    from bwpm_amd import adapter

    s = _structure(STACK)
    a, b = adapter.prepare(s, [LAMBDA_UM]), adapter.prepare(s, [LAMBDA_UM])
    assert a.content_hash == b.content_hash
    assert adapter.prepare(_structure(STACK[:-1] + [1.5 + 0j]), [LAMBDA_UM]).content_hash != a.content_hash
    inc = adapter.plane_wave(adapter.front_port(a, LAMBDA_UM), 1, 0, (1.0, 0.5))
    one = adapter.apply(a, inc, 3, precision=precision)
    two = adapter.apply(b, inc, 3, precision=precision)
    assert np.array_equal(one["back"].field, two["back"].field)
    assert np.array_equal(one["front"].field, two["front"].field)


def test_port_grid_is_centred() -> None:
    from bwpm_amd import adapter

    handle = adapter.prepare(_structure(STACK), [LAMBDA_UM])
    port = adapter.front_port(handle, LAMBDA_UM)
    y = (np.arange(32) - 16) * STEP
    beam = np.exp(-((y[:, None] / (3 * STEP)) ** 2 + (y[None, :] / (3 * STEP)) ** 2))
    field = np.stack([beam, 0.3j * beam], axis=-1)
    a = adapter.apply(handle, adapter.PortField(port, field), 3)
    b = adapter.apply(handle, adapter.PortField(port, np.roll(field, (2, -3), axis=(0, 1))), 3)
    assert np.max(np.abs(np.roll(a["back"].field, (2, -3), axis=(0, 1)) - b["back"].field)) < 1e-13


def test_backend_selection_and_estimate() -> None:
    from bwpm_amd import adapter, fast

    assert adapter.describe()["backends"] == ["gpu"]
    handle = adapter.prepare(_structure(STACK), [LAMBDA_UM])
    est = adapter.estimate(handle, iterations=3)
    assert est["seconds"] > 0 and est["memory_bytes"] > 0
    d = load_golden("fpwpm2d_interface")
    with pytest.raises(RuntimeError):
        fast.run_2d(_prop2d(d), d["inputs"]["field"], iterations=0, backend="cpu")
    with pytest.raises(ValueError):
        fast.run_2d(_prop2d(d), d["inputs"]["field"], iterations=0, backend="quantum")


def test_reference_solver_name_equals_the_gpu_solver_bitwise() -> None:
    """`_core.Solver(..., threads, fluctuation=...)` (the reference's non-CUDA branch,
    src/bwpm/fast.py:98) gives the bits of `_core.CudaSolver(..., device, fluctuation=...)`."""
    from bwpm_amd import _core, workloads

    p = workloads.tilted_3d(n=64, nz=24, iterations=2)
    a = _core.Solver(*p.solver_args(), 8, fluctuation=None)
    b = _core.CudaSolver(*p.solver_args(), 0, fluctuation=None)
    try:
        ra, rb = a.run(p.field, 2, [0, 12]), b.run(p.field, 2, [0, 12])
    finally:
        a.close()
        b.close()
    for key in ("transmitted", "reflected"):
        assert np.array_equal(ra[key], rb[key])
    for q in (0, 12):
        assert np.array_equal(ra["forward"][q], rb["forward"][q])
    assert ra["stored_values"] == rb["stored_values"]
