"""Index fluctuations generated on the device (bwpm_solver_create_synth, bwpm_queue_submit_synth).

The device's random stream is its own, so the generated volume is checked by its statistics: real
values, standard deviation σ, and the correlation of filtered noise. The filter exp(−|q|²l²/4)
gives a covariance ∝ exp(−r²/(2l²)), i.e. 0.607 at a lag of one correlation length. The solve on a
generated volume is checked against the oracle run on the same δn downloaded from the device.
Results are deterministic per seed, and a uniform label volume equals the same labels passed as an
array.
"""

from __future__ import annotations

import numpy as np
import pytest

from bwpm_amd import _core, serve, workloads
from bwpm_amd.specs import FluctuationSynthesis, Geometry, Layer, Uniform
from oracle import bwpm_oracle as O

pytestmark = pytest.mark.gpu


def _solver(p: workloads.Problem, precision: str = "fp64", labels=None) -> _core.GpuSolver:
    return _core.GpuSolver(p.ny, p.nx, p.dy, p.dx, p.dz, p.k0, p.indices, p.labels if labels is None else labels,
                           p.index_before, p.mask, fluctuation=p.fluctuation, precision=precision)


def _lag_corr(d: np.ndarray, axis: int, lag: int) -> float:
    a = d - d.mean()
    return float(np.mean(a * np.roll(a, -lag, axis=axis)) / np.mean(a * a))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_generated_fluctuation_statistics(precision: str) -> None:
    p = workloads.born_3d_synth(n=64, nz=64, iterations=1, sigma=0.01, corr_um=0.2, seed=11)
    s = _solver(p, precision)
    d = s.fluctuation()
    s.close()
    assert d.shape == (64, 64, 64)
    re = d.real
    assert np.max(np.abs(d.imag)) < (1e-12 if precision == "fp64" else 1e-6) * 0.01
    assert abs(re.std() / 0.01 - 1) < (1e-9 if precision == "fp64" else 1e-5)
    assert abs(re.mean()) < 0.2 * 0.01
    lag = 2  # corr 0.2 µm on a 0.1 µm grid; about 2000 independent samples, standard error ≈ 0.02
    for axis in (0, 1, 2):
        assert abs(_lag_corr(re, axis, lag) - np.exp(-0.5)) < 0.08, axis
        assert abs(_lag_corr(re, axis, 3 * lag)) < 0.12, axis


def test_generation_is_deterministic_per_seed() -> None:
    p = workloads.born_3d_synth(n=32, nz=12, iterations=1, sigma=0.02, corr_um=0.2, seed=5)
    a = _solver(p)
    b = _solver(p)
    q = workloads.born_3d_synth(n=32, nz=12, iterations=1, sigma=0.02, corr_um=0.2, seed=6)
    c = _solver(q)
    try:
        assert np.array_equal(a.fluctuation(), b.fluctuation())
        assert not np.array_equal(a.fluctuation(), c.fluctuation())
        ra, rb = a.run(p.field, 1), b.run(p.field, 1)
        assert np.array_equal(ra["transmitted"], rb["transmitted"])
        assert np.array_equal(ra["reflected"], rb["reflected"])
    finally:
        for s in (a, b, c):
            s.close()


@pytest.mark.parametrize("iterations", [1, 2])
def test_solve_on_generated_volume_equals_oracle(iterations: int) -> None:
    p = workloads.born_3d_synth(n=32, nz=10, iterations=iterations, sigma=0.02, corr_um=0.25, seed=21)
    s = _solver(p)
    try:
        out = s.run(p.field, iterations, [0, 5, 10])
        delta = s.fluctuation()
    finally:
        s.close()
    labels = np.zeros((p.nz, p.ny, p.nx), dtype=np.int32)
    ref = O.Solver(p.ny, p.nx, p.dy, p.dx, p.dz, p.k0, p.indices, labels, p.index_before, None,
                   fluctuation=delta).run(p.field, iterations, [0, 5, 10])
    scale = float(np.max(np.abs(p.field)))
    pairs = [(out["transmitted"], ref["transmitted"]), (out["reflected"], ref["reflected"])]
    for q in (0, 5, 10):
        pairs += [(out["forward"][q], ref["forward"][q]), (out["backward"][q], ref["backward"][q])]
    for got, want in pairs:
        assert np.max(np.abs(got - want)) < 1e-10 * scale


def test_uniform_volume_equals_label_array() -> None:
    p = workloads.born_3d_synth(n=32, nz=8, iterations=1, sigma=0.01, corr_um=0.2, seed=9)
    a = _solver(p)
    b = _solver(p, labels=np.zeros((p.nz, p.ny, p.nx), dtype=np.int32))
    try:
        ra, rb = a.run(p.field, 1), b.run(p.field, 1)
        assert np.array_equal(ra["transmitted"], rb["transmitted"])
        assert np.array_equal(ra["reflected"], rb["reflected"])
        assert np.array_equal(a.fluctuation(), b.fluctuation())
    finally:
        a.close()
        b.close()


def test_generated_requests_on_the_queue_equal_the_solves_alone() -> None:
    reqs = workloads.uneven_batch(count=4, sizes=(32, 64), nz_range=(6, 14))
    with serve.Server(slots=4, chunk=8) as srv:
        tickets = [srv.submit(p) for p in reqs]
        results = [srv.result(t, timeout=120) for t in tickets]
    for p, res in zip(reqs, results):
        s = _solver(p)
        try:
            alone = s.run(p.field, p.iterations)
        finally:
            s.close()
        assert np.array_equal(res["transmitted"], alone["transmitted"])
        assert np.array_equal(res["reflected"], alone["reflected"])


@pytest.mark.parametrize("precision", ["fp64", "mixed", "bfp"])
def test_geometry_volumes_equal_label_arrays(precision: str) -> None:
    cases = [(workloads.tilted_3d(n=64, nz=30, iterations=2), workloads.tilted_3d(n=64, nz=30, iterations=2,
                                                                                  geometry=True)),
             (serve.planar_3d(96, 30, 2, 1.0, 1.0, serve.SILVER, layer=(10, 13), geometry=False),
              serve.planar_3d(96, 30, 2, 1.0, 1.0, serve.SILVER, layer=(10, 13), geometry=True)),
             # an untilted plane: one label per slab, kept without a label volume
             (serve.planar_3d(128, 24, 2, 0.9, 1.0, 1.5, geometry=False),
              serve.planar_3d(128, 24, 2, 0.9, 1.0, 1.5, geometry=True))]
    for a, g in cases:
        sa, sg = _solver(a, precision), _solver(g, precision)
        try:
            ra, rg = sa.run(a.field, a.iterations, [7]), sg.run(g.field, g.iterations, [7])
        finally:
            sa.close()
            sg.close()
        assert ra["stored_values"] == rg["stored_values"] > 0
        for k in ("transmitted", "reflected"):
            assert np.array_equal(ra[k], rg[k])
        assert np.array_equal(ra["forward"][7], rg["forward"][7])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_composite_geometry_equals_its_label_array(precision: str) -> None:
    """A cylinder, a plane tilted in x and y over it and a layer: slabs where every primitive is
    uniform, slabs the plane or the cylinder cuts, and runs of equal slabs. The solver's labels made
    slab by slab must equal the materialised volume (bwpm_geometry_labels) bit for bit."""
    from bwpm_amd.specs import Above, Cylinder

    base = workloads.tilted_3d(n=64, nz=40, iterations=2, geometry=True)
    dz, dx = base.dz, base.dx
    base.indices = [1.0 + 0j, 1.46 + 0j, 1.3 + 0.01j]
    geo = Geometry(40, (Cylinder(2, 12 * dx, 8 * dz, 20 * dz, 3 * dx, -2 * dx),
                        Above(1, 24 * dz, 0.05, -0.08), Layer(2, 33 * dz, 35 * dz)))
    lab = _core.geometry_labels(geo, base.ny, base.nx, base.dy, base.dx, base.dz, 3)
    assert len(np.unique(lab)) == 3
    outs = []
    for labels in (geo, lab):
        s = _core.GpuSolver(base.ny, base.nx, base.dy, base.dx, base.dz, base.k0, base.indices, labels,
                            base.index_before, None, precision=precision)
        try:
            outs.append(s.run(base.field, 2, [10, 30]))
        finally:
            s.close()
    for k in ("transmitted", "reflected"):
        assert np.array_equal(outs[0][k], outs[1][k])
    for q in (10, 30):
        assert np.array_equal(outs[0]["forward"][q], outs[1]["forward"][q])
    assert outs[0]["stored_values"] == outs[1]["stored_values"] > 0


def test_geometry_with_generated_fluctuation_equals_the_label_array() -> None:
    geo = Geometry(12, (Layer(1, 4e-4, 8e-4),))
    p = workloads.born_3d_synth(n=32, nz=12, iterations=1, sigma=0.02, corr_um=0.2, seed=4)
    p.indices = [1.5 + 0j, 1.6 + 0j]
    lab = _core.geometry_labels(geo, p.ny, p.nx, p.dy, p.dx, p.dz, 2)
    assert 0 < lab.sum() < lab.size
    outs = []
    for labels in (geo, lab):
        s = _core.GpuSolver(p.ny, p.nx, p.dy, p.dx, p.dz, p.k0, p.indices, labels, p.index_before, None,
                            fluctuation=p.fluctuation)
        try:
            outs.append((s.run(p.field, 1), s.fluctuation()))
        finally:
            s.close()
    assert np.array_equal(outs[0][0]["transmitted"], outs[1][0]["transmitted"])
    assert np.array_equal(outs[0][0]["reflected"], outs[1][0]["reflected"])
    assert np.array_equal(outs[0][1], outs[1][1])


def test_geometry_requests_on_the_queue_equal_the_solves_alone() -> None:
    reqs = [workloads.tilted_3d(n=64, nz=24 + 4 * i, iterations=1 + i % 2, geometry=True, tilt_deg=5.0 + i)
            for i in range(4)]
    reqs.append(serve.planar_3d(64, 20, 1, 1.0, 1.0, serve.SILVER, layer=(8, 11)))
    with serve.Server(slots=3, chunk=8) as srv:
        tickets = [srv.submit(p) for p in reqs]
        results = [srv.result(t, timeout=120) for t in tickets]
    for p, res in zip(reqs, results):
        s = _solver(p)
        try:
            alone = s.run(p.field, p.iterations)
        finally:
            s.close()
        assert np.array_equal(res["transmitted"], alone["transmitted"])
        assert np.array_equal(res["reflected"], alone["reflected"])


def test_synthesis_argument_errors() -> None:
    p = workloads.born_3d_synth(n=16, nz=4, iterations=1)
    with pytest.raises(ValueError):
        _core.GpuSolver(p.ny, p.nx, p.dy, p.dx, p.dz, p.k0, p.indices, p.labels, p.index_before, None,
                        fluctuation=FluctuationSynthesis(-1.0, 1e-3, 1))
    with pytest.raises(ValueError):
        _core.GpuSolver(p.ny, p.nx, p.dy, p.dx, p.dz, p.k0, p.indices, Uniform(4, 3), p.index_before, None,
                        fluctuation=p.fluctuation)
    lab = np.zeros((4, 16, 16), dtype=np.int32)
    lab[2, :, :8] = 1
    with pytest.raises(ValueError, match="host"):
        _core.GpuSolver(p.ny, p.nx, p.dy, p.dx, p.dz, p.k0, [1.5, 1.0], lab, p.index_before, None,
                        fluctuation=p.fluctuation)
