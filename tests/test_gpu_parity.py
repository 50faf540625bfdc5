"""GPU parity of the slab-stepping path through the C ABI, against the reference's golden fixtures
and the FP64 oracle.

Tolerances: FP64 mode as the reference's own GPU tests (rel 1e-10, abs 1e-11·max,
tests/test_core_cuda.py:34-36); FP32 mode by relative field L2 ≤ 1e-5 (the bulk-precision bar).
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden, rel_l2
from helpers import fixture_2d_args, fixture_3d_args, random_volume
from oracle import bwpm_oracle as O

pytestmark = pytest.mark.gpu

L2_FP32 = 1e-5
# complex64 streaming planes with FP64 arithmetic: only the storage rounding of two planes per step
# (weak reflected planes of the random silver volumes see ~1.3e-6); block-floating-point planes
# round 8 to 16 times finer
L2 = {"mixed": 5e-6, "bfp": 1e-6, "fp32": L2_FP32}


def _core():
    from bwpm_amd import _core

    assert _core.gpu_device_count() > 0, "no GPU visible"
    return _core


def _close64(a: np.ndarray, b: np.ndarray) -> None:
    assert a.shape == b.shape
    atol = 1e-11 * float(np.max(np.abs(b)) or 1.0)
    err = float(np.max(np.abs(a - b)))
    assert np.allclose(a, b, rtol=1e-10, atol=atol), err


def _solver(args: dict, precision: str):
    c = _core()
    return c.GpuSolver(args["ny"], args["nx"], args["dy"], args["dx"], args["dz"], args["k0"],
                       args["indices"], args["labels"], args["index_before"], args.get("mask"),
                       fluctuation=args.get("fluctuation"), precision=precision)


def test_interface_matrices_fixture() -> None:
    c = _core()
    d = load_golden("interface_matrices")
    kx, ky = d["inputs"]["k_x_per_mm"], d["inputs"]["k_y_per_mm"]
    k0 = d["parameters"]["k0_per_mm"]
    tol = d["tolerance"]
    for j, (n1, n2) in enumerate(d["inputs"]["index_pairs"]):
        s = n1.real
        t = c.transmission_matrix(kx * s, ky * s, n1, n2, k0)
        r = c.reflection_matrix(kx * s, ky * s, n1, n2, k0)
        assert np.allclose(t, d["expected"]["T"][j], rtol=tol["rel"], atol=tol["abs"])
        assert np.allclose(r, d["expected"]["R"][j], rtol=tol["rel"], atol=tol["abs"])


@pytest.mark.parametrize(("n1", "n2"), [(1.0, 1.5), (1.5, 1.0), (1.0, 0.213 + 6.697j), (1.3, 1.3)])
def test_interface_on_fft_grid_with_grazing_points(n1: complex, n2: complex) -> None:
    """k⊥ = 0, propagating, exactly grazing and evanescent points (test_core_interface.py:42-56)."""
    c = _core()
    k0 = 2 * np.pi / 0.5e-3
    k = np.arange(-8, 8) * k0 / 4
    kx, ky = np.meshgrid(k, k)
    for fn, ref in ((c.transmission_matrix, O.transmission_matrix), (c.reflection_matrix, O.reflection_matrix)):
        a, b = fn(kx, ky, n1, n2, k0), ref(kx, ky, n1, n2, k0)
        assert a.shape == (*kx.shape, 2, 2)
        assert np.all(np.isfinite(a))
        assert np.allclose(a, b, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize(("n1", "n2"), [(1.0, 1.5), (1.5, 1.0)])
def test_grazing_rule_is_independent_of_rounding(n1: complex, n2: complex) -> None:
    """Perturbing k⊥ by a few ulp at |k⊥| = n₁k₀ must not change the matrices
    (test_core_interface.py:68-85)."""
    c = _core()
    k0 = 2 * np.pi / 0.5e-3
    for angle in (0.0, 0.3, 1.1):
        k = n1 * k0
        for rel in (0.0, 3e-16, -3e-16, 1e-14):
            kx = np.array([k * (1 + rel) * np.cos(angle)])
            ky = np.array([k * (1 + rel) * np.sin(angle)])
            for fn, ref in ((c.transmission_matrix, O.transmission_matrix), (c.reflection_matrix, O.reflection_matrix)):
                assert np.allclose(fn(kx, ky, n1, n2, k0), ref(kx, ky, n1, n2, k0), rtol=1e-13, atol=1e-13)


def test_longitudinal_wavenumber_matches_oracle() -> None:
    c = _core()
    k0 = 2 * np.pi / 0.8e-3
    kp2 = np.linspace(0, 4 * k0 * k0, 101)
    for n in (1.0, 1.5, 0.13 + 4.0j, 2.1 + 0.01j):
        a = c.longitudinal_wavenumber(n, k0, kp2)
        b = O.longitudinal_wavenumber(n, k0, kp2)
        assert np.allclose(a, b, rtol=1e-14, atol=1e-12 * k0)
    assert isinstance(c.longitudinal_wavenumber(1.5, k0, 0.0), complex)


@pytest.mark.parametrize("precision", ["fp64", "mixed", "bfp", "fp32"])
def test_fpwpm2d_interface(precision: str) -> None:
    d = load_golden("fpwpm2d_interface")
    a = fixture_2d_args(d)
    nz = a["labels"].shape[0]
    out = _solver(a, precision).run(a["field"], 0, list(range(nz + 1)))
    planes = np.stack([out["forward"][p][:, 0] for p in range(nz + 1)])
    if precision == "fp64":
        _close64(planes, d["expected"]["planes"])
    else:
        assert rel_l2(planes, d["expected"]["planes"]) < L2[precision]
    assert out["backward"] == {} and out["stored_values"] == 0


@pytest.mark.parametrize("precision", ["fp64", "mixed", "bfp", "fp32"])
def test_bwpm2d_tilted_silver(precision: str) -> None:
    d = load_golden("bwpm2d_tilted_silver")
    a = fixture_2d_args(d)
    nz = a["labels"].shape[0]
    out = _solver(a, precision).run(a["field"], 2, list(range(nz + 1)))
    fwd = np.stack([out["forward"][p][:, 0] for p in range(nz + 1)])
    bwd = np.stack([out["backward"][p][:, 0] for p in range(nz + 1)])
    p = d["parameters"]
    last = complex(d["inputs"]["indices"][int(d["inputs"]["labels"][-1][p["ny"] // 2])])
    res = O.energy_residual_2d(d["inputs"]["field"], [t[:, 0] for t in out["transmitted_iter"]],
                               [r[:, 0] for r in out["reflected_iter"]], p["ny"], p["dy_mm"], a["k0"], 1.0, last)
    if precision == "fp64":
        _close64(fwd, d["expected"]["forward"])
        _close64(bwd, d["expected"]["backward"])
        assert np.allclose(res, d["expected"]["energy_residual"], rtol=1e-10, atol=1e-12)
    else:
        assert rel_l2(fwd, d["expected"]["forward"]) < L2[precision]
        assert rel_l2(bwd, d["expected"]["backward"]) < L2[precision]
        assert np.allclose(res, d["expected"]["energy_residual"], atol=1e-4)


@pytest.mark.parametrize("precision", ["fp64", "mixed", "bfp", "fp32"])
def test_bwpm2d_slab_exact(precision: str) -> None:
    d = load_golden("bwpm2d_slab_exact")
    a = fixture_2d_args(d)
    out = _solver(a, precision).run(a["field"], 40)
    t, r = out["transmitted"][:, 0], out["reflected"][:, 0]
    if precision == "fp64":
        _close64(t, d["expected"]["transmitted"])
        _close64(r, d["expected"]["reflected"])
        rr = complex(np.mean(r[:, 0] / d["inputs"]["field"][:, 0]))
        assert abs(rr - d["reference"]["r_te_airy"]) < d["reference_tolerance"]["abs"]
    else:
        assert rel_l2(t, d["expected"]["transmitted"]) < L2[precision]
        assert rel_l2(r, d["expected"]["reflected"]) < L2[precision]


@pytest.mark.parametrize("precision", ["fp64", "mixed", "bfp", "fp32"])
def test_bwpm3d_tilted(precision: str) -> None:
    d = load_golden("bwpm3d_tilted")
    a = fixture_3d_args(d)
    nz = a["labels"].shape[0]
    keep = (0, 10, nz)
    out = _solver(a, precision).run(a["field"], 2, keep)
    fwd = np.stack([out["forward"][q] for q in keep])
    bwd = np.stack([out["backward"][q] for q in keep])
    e = d["expected"]
    if precision == "fp64":
        _close64(out["transmitted"], e["transmitted"])
        _close64(out["reflected"], e["reflected"])
        _close64(fwd, e["forward_planes"])
        _close64(bwd, e["backward_planes"])
    else:
        for got, want in ((out["transmitted"], e["transmitted"]), (out["reflected"], e["reflected"]),
                          (fwd, e["forward_planes"]), (bwd, e["backward_planes"])):
            assert rel_l2(got, want) < L2[precision]


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("precision", ["fp64", "mixed", "bfp", "fp32"])
def test_random_volumes_equal_oracle(seed: int, precision: str) -> None:
    """Random volumes with four materials incl. silver, odd and even sizes, dx ≠ dy (grazing
    points), 2 iterations; the reference's core-vs-Python test (test_core_solver.py:99-127)."""
    a = random_volume(seed)
    nz = a["labels"].shape[0]
    keep = (0, 5, nz)
    ref = O.Solver(a["ny"], a["nx"], a["dy"], a["dx"], a["dz"], a["k0"], a["indices"], a["labels"],
                   a["index_before"], a["mask"]).run(a["field"], 2, keep)
    out = _solver(a, precision).run(a["field"], 2, keep)
    assert out["stored_values"] == ref["stored_values"]
    pairs = [(out["transmitted"], ref["transmitted"]), (out["reflected"], ref["reflected"])]
    for q in keep:
        pairs += [(out["forward"][q], ref["forward"][q]), (out["backward"][q], ref["backward"][q])]
    scale = float(np.max(np.abs(a["field"])))
    for got, want in pairs:
        if precision == "fp64":
            assert np.max(np.abs(got - want)) < 1e-10 * scale
        else:
            assert rel_l2(got, want) < L2[precision]


# ---- Born scattering at index fluctuations -------------------------------------------------------

@pytest.mark.parametrize("precision", ["fp64", "mixed", "bfp", "fp32"])
def test_fpwpm2d_fluctuation(precision: str) -> None:
    d = load_golden("fpwpm2d_fluctuation")
    a = fixture_2d_args(d)
    nz = a["labels"].shape[0]
    out = _solver(a, precision).run(a["field"], 0, list(range(nz + 1)))
    planes = np.stack([out["forward"][p][:, 0] for p in range(nz + 1)])
    if precision == "fp64":
        _close64(planes, d["expected"]["planes"])
    else:
        assert rel_l2(planes, d["expected"]["planes"]) < L2[precision]
    assert out["gmres_iterations"] > 0


@pytest.mark.parametrize("precision", ["fp64", "mixed", "bfp", "fp32"])
def test_bwpm3d_fluctuation(precision: str) -> None:
    d = load_golden("bwpm3d_fluctuation")
    a = fixture_3d_args(d)
    nz = a["labels"].shape[0]
    keep = (0, 8, nz)
    out = _solver(a, precision).run(a["field"], 1, keep)
    e = d["expected"]
    pairs = [(out["transmitted"], e["transmitted"]), (out["reflected"], e["reflected"]),
             (np.stack([out["forward"][q] for q in keep]), e["forward_planes"]),
             (np.stack([out["backward"][q] for q in keep]), e["backward_planes"])]
    for got, want in pairs:
        if precision == "fp64":
            _close64(got, want)
        else:
            assert rel_l2(got, want) < L2[precision]


def _fluct_volume(seed: int, back: bool):
    from helpers import k0_of, mask_2d

    rng = np.random.default_rng(seed)
    if back:  # tests/test_core_solver.py:224-257 construction
        nx, ny, nz = (12, 10, 14) if seed % 2 else (8, 16, 11)
        dx, dz, k0 = 0.05e-3, 0.04e-3, k0_of(0.8)
        labels = np.zeros((nz, ny, nx), dtype=np.int32)
        labels[nz - 3:] = 1
        fl = np.zeros((nz, ny, nx), complex)
        fl[1:nz - 4] = 0.02 * rng.normal(size=(nz - 5, ny, nx)) + 1e-3j
        field = rng.normal(size=(ny, nx, 2)) + 1j * rng.normal(size=(ny, nx, 2))
        mask = mask_2d(ny, nx, dx, dx, 2 * dx) if seed == 4 else None
        return dict(ny=ny, nx=nx, dy=dx, dx=dx, dz=dz, k0=k0, indices=[1.38, 1.5 + 0.002j], labels=labels,
                    index_before=1.38, mask=mask, fluctuation=fl, field=field), 3
    # tests/test_core_solver.py:176-203 construction: a mixed slab without δn, grazing points
    nx, ny, nz = (15, 12, 9) if seed % 2 else (8, 16, 11)
    dx, dy, dz, k0 = 0.04e-3, 0.05e-3, 0.03e-3, k0_of(0.6)
    labels = rng.integers(0, 2, size=(nz, 1, 1)) * np.ones((1, ny, nx), dtype=np.int64)
    labels[nz // 2] = rng.integers(0, 3, size=(ny, nx))
    fl = 0.03 * rng.normal(size=(nz, ny, nx)) + 0.002j * rng.normal(size=(nz, ny, nx))
    fl[nz // 2] = 0.0
    field = rng.normal(size=(ny, nx, 2)) + 1j * rng.normal(size=(ny, nx, 2))
    mask = mask_2d(ny, nx, dy, dx, 2 * dx) if seed == 2 else None
    return dict(ny=ny, nx=nx, dy=dy, dx=dx, dz=dz, k0=k0, indices=[1.33, 1.5, 2.1 + 0.01j], labels=labels,
                index_before=1.33, mask=mask, fluctuation=fl, field=field), 1


@pytest.mark.parametrize(("seed", "back"), [(1, False), (2, False), (3, True), (4, True)])
@pytest.mark.parametrize("precision", ["fp64", "mixed", "bfp", "fp32"])
def test_fluctuations_equal_oracle(seed: int, back: bool, precision: str) -> None:
    a, its = _fluct_volume(seed, back)
    nz = a["labels"].shape[0]
    keep = (0, 5, nz)
    ref = O.Solver(a["ny"], a["nx"], a["dy"], a["dx"], a["dz"], a["k0"], a["indices"], a["labels"],
                   a["index_before"], a["mask"], fluctuation=a["fluctuation"]).run(a["field"], its, keep)
    out = _solver(a, precision).run(a["field"], its, keep)
    pairs = [(out["transmitted"], ref["transmitted"]), (out["reflected"], ref["reflected"])]
    for q in keep:
        pairs += [(out["forward"][q], ref["forward"][q]), (out["backward"][q], ref["backward"][q])]
    scale = float(np.max(np.abs(a["field"])))
    for got, want in pairs:
        if precision == "fp64":
            assert np.max(np.abs(got - want)) < 1e-10 * scale
        else:
            assert rel_l2(got, want) < L2[precision]


def test_born_step_conserves_flux_and_absorbs_exactly() -> None:
    """K20 invariants on the GPU (tests/test_stable_born.py:66-121): strong real δn conserves the
    flux; uniform δn = iκ attenuates as e^{-2k0κL}."""
    from helpers import k0_of

    def flux(f, n0, k0, d):
        n = f.shape[0]
        kx = O.fftfreq(n, d)[None, :]
        ky = O.fftfreq(n, d)[:, None]
        k2 = (n0 * k0) ** 2
        prop = kx ** 2 + ky ** 2 < k2 * (1 - 1e-12)
        kz = np.sqrt(np.where(prop, k2 - kx ** 2 - ky ** 2, 1.0))
        sp = np.fft.fft2(f, axes=(0, 1))
        ez = -(kx * sp[..., 0] + ky * sp[..., 1]) / kz
        return float(np.sum(np.where(prop, kz * (np.sum(np.abs(sp) ** 2, -1) + np.abs(ez) ** 2), 0.0)))

    k0 = k0_of(0.8)
    n, nz, h = 24, 12, 0.1e-3
    f = np.zeros((n, n, 2), complex)
    f[..., 0] = 1.0
    delta = 0.1 * np.random.default_rng(7).normal(size=(nz, n, n))
    s = _core().GpuSolver(n, n, h, h, h, k0, [1.38], np.zeros((nz, n, n), np.int32), 1.38, fluctuation=delta)
    out = s.run(f, 0)["transmitted"]
    assert abs(flux(out, 1.38, k0, h) / flux(f, 1.38, k0, h) - 1) < 1e-8
    n, nz, h, kappa = 8, 100, 0.05e-3, 1e-3
    f = np.zeros((n, n, 2), complex)
    f[..., 0] = 1.0
    s = _core().GpuSolver(n, n, h, h, h, k0, [1.38], np.zeros((nz, n, n), np.int32), 1.38,
                          fluctuation=np.full((nz, n, n), 1j * kappa))
    out = s.run(f, 0)["transmitted"]
    expected = np.exp(-2 * k0 * kappa * nz * h)
    assert abs(flux(out, 1.38, k0, h) / flux(f, 1.38, k0, h) / expected - 1) < 1e-6


def test_two_hosts_in_a_slab_is_a_value_error() -> None:
    labels = np.zeros((2, 4, 4), np.int32)
    labels[0, :, :2] = 1
    fl = np.zeros((2, 4, 4))
    fl[0] = 0.01
    with pytest.raises(ValueError, match="host"):
        _core().GpuSolver(4, 4, 1e-4, 1e-4, 1e-4, 1e4, [1.0, 1.5], labels, 1.0, fluctuation=fl)


# ---- the device path against the reference's own C++ core (oracle/_ref, built from its sources) --

def _reference_core():
    import importlib.util

    from conftest import ROOT

    ref_dir = ROOT / "oracle" / "_ref"
    so = sorted(ref_dir.glob("_core*.so")) if ref_dir.exists() else []
    if not so:
        pytest.skip("reference core not built (oracle/build_ref.sh)")
    spec = importlib.util.spec_from_file_location("_core", so[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("pol", ["x", "y"])
def test_device_equals_reference_core_tilted_256(pol: str) -> None:
    """The FP64 register-kernel path (interface steps with several label pairs, reflections, two
    iterations) against the reference's C++ core on the 256² tilted geometry of config 2, 64 slabs:
    agreement to 1e-12 of the field's maximum, and the same number of stored interface values."""
    from bwpm_amd import _core, workloads

    R = _reference_core()
    p = workloads.tilted_3d(n=256, nz=64, iterations=2, polarisation=pol)
    s = _core.GpuSolver(*p.solver_args(), precision="fp64")
    try:
        got = s.run(p.field, p.iterations, [0, 32, 64])
    finally:
        s.close()
    ref = R.Solver(p.ny, p.nx, p.dy, p.dx, p.dz, p.k0, np.asarray(p.indices, complex), p.labels.astype(np.int32),
                   complex(p.index_before), None, 16).run(p.field, p.iterations, [0, 32, 64])
    scale = float(np.max(np.abs(p.field)))
    for key in ("transmitted", "reflected"):
        assert np.max(np.abs(got[key] - ref[key])) < 1e-12 * scale, key
    for q in (0, 32, 64):
        assert np.max(np.abs(got["forward"][q] - ref["forward"][q])) < 1e-12 * scale
        assert np.max(np.abs(got["backward"][q] - ref["backward"][q])) < 1e-12 * scale
    assert got["stored_values"] == ref["stored_values"]


@pytest.mark.parametrize("restart", [2, 3])
def test_born_gmres_restarts_match_the_oracle(restart: int, monkeypatch) -> None:
    """A GMRES restart length of 2 or 3 (BWPM_GMRES_RESTART) makes every Born step end cycles
    without converging, so the restart path of the device cycle (a full cycle, the update of x,
    a new r₀) runs; the field still matches the oracle (restart 30) to 1e-10."""
    from bwpm_amd import _core

    monkeypatch.setenv("BWPM_GMRES_RESTART", str(restart))
    monkeypatch.setenv("BWPM_BORN_SOLVER", "gmres")
    rng = np.random.default_rng(11)
    ny = nx = 32
    nz = 8
    dx, dz, lam = 0.1e-3, 0.15e-3, 0.8
    k0 = 2 * np.pi / (lam * 1e-3)
    labels = np.zeros((nz, ny, nx), dtype=np.int32)
    fl = np.zeros((nz, ny, nx), complex)
    fl[1:nz - 1] = 0.05 * rng.normal(size=(nz - 2, ny, nx))
    x = (np.arange(nx) - nx / 2) * dx
    field = np.zeros((ny, nx, 2), complex)
    field[..., 0] = np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / (1.0e-3) ** 2)
    indices = np.array([1.5])
    s = _core.GpuSolver(ny, nx, dx, dx, dz, k0, indices, labels, 1.5 + 0j, None, fluctuation=fl, precision="fp64")
    try:
        got = s.run(field, 2)
    finally:
        s.close()
    ref = O.Solver(ny, nx, dx, dx, dz, k0, indices, labels, 1.5 + 0j, None, fluctuation=fl).run(field, 2)
    scale = float(np.max(np.abs(field)))
    assert got["gmres_iterations"] > (nz - 2) * 2 * 2 * restart  # more than one cycle per Born step
    for key in ("transmitted", "reflected"):
        assert np.max(np.abs(got[key] - ref[key])) < 1e-10 * scale, key


@pytest.mark.parametrize("solver", ["gmres", "default"])
def test_born_solvers_agree_with_the_oracle(solver: str, monkeypatch) -> None:
    """The Cayley system solved by GMRES alone (BWPM_BORN_SOLVER=gmres) and by the default
    (Richardson steps, GMRES only where they stall) on weak, strong and very strong fluctuations
    (σ = 0.02, 0.5 and 2, the last beyond Richardson's contraction, so the default falls back to
    GMRES): both match the oracle to 1e-10 of the field's maximum."""
    from bwpm_amd import _core

    if solver == "gmres":
        monkeypatch.setenv("BWPM_BORN_SOLVER", "gmres")
    for sigma in (0.02, 0.5, 2.0):
        rng = np.random.default_rng(5)
        ny = nx = 32
        nz = 6
        dx, dz, lam = 0.1e-3, 0.15e-3, 0.8
        k0 = 2 * np.pi / (lam * 1e-3)
        labels = np.zeros((nz, ny, nx), dtype=np.int32)
        fl = np.zeros((nz, ny, nx), complex)
        fl[1:nz - 1] = sigma * rng.normal(size=(nz - 2, ny, nx))
        x = (np.arange(nx) - nx / 2) * dx
        field = np.zeros((ny, nx, 2), complex)
        field[..., 0] = np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / (1.0e-3) ** 2)
        s = _core.GpuSolver(ny, nx, dx, dx, dz, k0, np.array([1.5]), labels, 1.5 + 0j, None, fluctuation=fl,
                            precision="fp64")
        try:
            got = s.run(field, 2)
        finally:
            s.close()
        ref = O.Solver(ny, nx, dx, dx, dz, k0, [1.5], labels, 1.5 + 0j, None, fluctuation=fl).run(field, 2)
        for key in ("transmitted", "reflected"):
            scale = max(float(np.max(np.abs(field))), float(np.max(np.abs(ref[key]))))
            assert np.max(np.abs(got[key] - ref[key])) < 1e-10 * scale, (sigma, key)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize(("ny", "nx"), [(192, 192), (192, 384)])
def test_born_on_three_times_power_of_two_grids(precision: str, ny: int, nx: int) -> None:
    """Born slabs on grids of 3·2^k points, which run the register 2D transforms and the fused
    operator with a radix-3 or radix-6 pass (both components per thread), against the oracle over
    two iterations: FP64 to 1e-10 of the field's maximum, FP32 to 2e-5."""
    from bwpm_amd import _core

    rng = np.random.default_rng(11)
    nz = 4
    dx, dz, lam = 0.1e-3, 0.15e-3, 0.8
    k0 = 2 * np.pi / (lam * 1e-3)
    labels = np.zeros((nz, ny, nx), dtype=np.int32)
    fl = np.zeros((nz, ny, nx), complex)
    fl[1:nz - 1] = 0.02 * rng.normal(size=(nz - 2, ny, nx))
    x = (np.arange(nx) - nx / 2) * dx
    y = (np.arange(ny) - ny / 2) * dx
    field = np.zeros((ny, nx, 2), complex)
    field[..., 0] = np.exp(-(x[None, :] ** 2 + y[:, None] ** 2) / (2.0e-3) ** 2)
    field[..., 1] = 0.3j * field[..., 0]
    print("born 3·2^k", precision, ny, nx, flush=True)
    s = _core.GpuSolver(ny, nx, dx, dx, dz, k0, np.array([1.5]), labels, 1.5 + 0j, None, fluctuation=fl,
                        precision=precision)
    try:
        got = s.run(field, 2)
    finally:
        s.close()
    ref = O.Solver(ny, nx, dx, dx, dz, k0, [1.5], labels, 1.5 + 0j, None, fluctuation=fl).run(field, 2)
    tol = 1e-10 if precision == "fp64" else 2e-5
    for key in ("transmitted", "reflected"):
        scale = max(float(np.max(np.abs(field))), float(np.max(np.abs(ref[key]))))
        assert np.max(np.abs(got[key] - ref[key])) < tol * scale, key


def test_device_equals_reference_core_born() -> None:
    """Born slabs (δn, device GMRES) with the back-scattering paths over three iterations against
    the reference's C++ core, 64² × 12 slabs: agreement to 1e-10 of the field's maximum (GMRES stops
    at a relative residual of 1e-12 in both)."""
    from bwpm_amd import _core

    R = _reference_core()
    rng = np.random.default_rng(7)
    ny = nx = 64
    nz = 12
    dx, dz, lam = 0.1e-3, 0.15e-3, 0.8
    k0 = 2 * np.pi / (lam * 1e-3)
    labels = np.zeros((nz, ny, nx), dtype=np.int32)
    labels[nz - 3:] = 1
    fl = np.zeros((nz, ny, nx), complex)
    fl[1:nz - 4] = 0.01 * rng.normal(size=(nz - 5, ny, nx)) + 1e-3j
    x = (np.arange(nx) - nx / 2) * dx
    field = np.zeros((ny, nx, 2), complex)
    field[..., 0] = np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / (1.5e-3) ** 2)
    indices = np.array([1.38, 1.5 + 0.002j])
    s = _core.GpuSolver(ny, nx, dx, dx, dz, k0, indices, labels, 1.38 + 0j, None, fluctuation=fl, precision="fp64")
    try:
        got = s.run(field, 3)
    finally:
        s.close()
    ref = R.Solver(ny, nx, dx, dx, dz, k0, indices, labels, 1.38 + 0j, None, 16, fluctuation=fl).run(field, 3)
    scale = float(np.max(np.abs(field)))
    for key in ("transmitted", "reflected"):
        assert np.max(np.abs(got[key] - ref[key])) < 1e-10 * scale, key
