"""GPU parity of the register-resident slab kernels (power-of-two grids) and of batched lanes.

Volumes mix a homogeneous run, a tilted glass interface, an absorbing layer and silver, with an
absorbing edge mask and band-limited random fields (the construction of the reference's
tests/test_core_solver.py:99-127), compared with the FP64 oracle; batched solves must equal the
same solves run alone bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_l2
from helpers import SILVER, band_limited_field, k0_of, mask_2d
from oracle import bwpm_oracle as O

pytestmark = pytest.mark.gpu


def layered_volume(ny: int, nx: int, nz: int, seed: int, mask: bool = True) -> dict:
    rng = np.random.default_rng(seed)
    lam = 0.6
    dy, dx, dz = lam * 1e-3 / 6, lam * 1e-3 / 5, lam * 1e-3 / 4
    k0 = k0_of(lam)
    y = (np.arange(ny) - ny // 2) * dy
    zc = (np.arange(nz) + 0.5) * dz
    labels = np.zeros((nz, ny, nx), dtype=np.int32)
    tilt = (zc[:, None] > 3 * dz + 0.3 * y[None, :])
    labels[:] = np.where(tilt, 1, 0)[:, :, None]
    labels[nz - 3] = 2
    labels[nz - 2:] = 3
    field = band_limited_field(rng, ny, nx, dy, dx, k0)
    return dict(ny=ny, nx=nx, dy=dy, dx=dx, dz=dz, k0=k0, indices=[1.0, 1.5, 2.1 + 0.01j, SILVER],
                labels=labels, index_before=1.0 + 0j, mask=mask_2d(ny, nx, dy, dx, 3 * dx) if mask else None,
                field=field)


def _solver(a: dict, precision: str):
    from bwpm_amd import _core

    return _core.GpuSolver(a["ny"], a["nx"], a["dy"], a["dx"], a["dz"], a["k0"], a["indices"], a["labels"],
                           a["index_before"], a["mask"], precision=precision)


@pytest.mark.parametrize("shape", [(64, 64), (128, 128), (256, 128), (512, 512), (512, 12), (12, 256), (4096, 1),
                                   (2048, 2), (3072, 1), (384, 96), (80, 120), (1000, 1), (2, 1024), (1, 4096),
                                   (4, 2048),
                                   # lines of 3·2^k and 9·2^k points (irregular register passes)
                                   (192, 192), (384, 384), (768, 64), (192, 1536), (1, 576), (1, 1152), (2, 3072),
                                   (1536, 4), (576, 24), (1152, 2)])
@pytest.mark.parametrize("precision", ["fp64", "mixed", "bfp", "fp32"])
def test_register_kernels_match_oracle(shape: tuple[int, int], precision: str) -> None:
    ny, nx = shape
    nz = 9
    a = layered_volume(ny, nx, nz, seed=ny + nx)
    keep = (0, 4, nz)
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
        elif np.max(np.abs(want)) < 1e-6 * scale:
            # the field is extinguished (behind the silver layer at some shapes it is ~1e-64 of the
            # input, below FP32's range): the FP32 result must be as small
            assert np.max(np.abs(got)) < 1e-6 * scale
        else:
            assert rel_l2(got, want) < (1e-6 if precision in ("mixed", "bfp") else 1e-5)


@pytest.mark.parametrize("shape", [(128, 128), (256, 128), (512, 512), (1024, 2), (2048, 2), (192, 192), (384, 384),
                                   (768, 64), (1, 576), (1, 1152), (192, 1536), (2, 3072)])
@pytest.mark.parametrize("precision", ["fp64", "mixed", "fp32"])
def test_alternative_radix_plans_match_oracle(shape: tuple[int, int], precision: str, monkeypatch) -> None:
    """BWPM_PLAN=1 runs every register transform on its length's alternative radix plan (the one the
    plan tuner times against the default): the same results as the oracle."""
    monkeypatch.setenv("BWPM_PLAN", "1")
    ny, nx = shape
    nz = 7
    a = layered_volume(ny, nx, nz, seed=3 * ny + nx)
    ref = O.Solver(a["ny"], a["nx"], a["dy"], a["dx"], a["dz"], a["k0"], a["indices"], a["labels"],
                   a["index_before"], a["mask"]).run(a["field"], 2)
    out = _solver(a, precision).run(a["field"], 2)
    scale = float(np.max(np.abs(a["field"])))
    for key in ("transmitted", "reflected"):
        got, want = out[key], ref[key]
        if precision == "fp64":
            assert np.max(np.abs(got - want)) < 1e-10 * scale, key
        elif np.max(np.abs(want)) < 1e-6 * scale:
            assert np.max(np.abs(got)) < 1e-6 * scale, key
        else:
            assert rel_l2(got, want) < (1e-6 if precision == "mixed" else 1e-5), key


def test_plan_tuner_times_both_plans() -> None:
    """Without BWPM_PLAN the tuner times both radix plans of a length once, before its first solve,
    and every later solve of that length takes the faster one: solves after the tuning give the same
    bits as each other."""
    import os

    from bwpm_amd import _core

    assert "BWPM_PLAN" not in os.environ
    a = layered_volume(384, 384, 6, seed=11)
    first = _solver(a, "mixed").run(a["field"], 1)
    plans = [p for p in _core.fft_plans() if p["n"] == 384 and p["precision"] == 48]
    assert {p["direction"] for p in plans} == {"col", "row"}
    for p in plans:
        assert p["us"][0] > 0 and p["us"][1] > 0
        assert p["plan"] == (1 if p["us"][1] < p["us"][0] else 0)
    again = _solver(a, "mixed").run(a["field"], 1)
    for key in ("transmitted", "reflected"):
        assert np.array_equal(first[key], again[key])


@pytest.mark.parametrize("precision", ["fp32", "mixed", "bfp", "fp64"])
def test_batch_equals_alone_bitwise(precision: str) -> None:
    """Solves of two grid sizes, uneven depths and different geometries: the batched lanes give
    each solve exactly its result alone."""
    from bwpm_amd import _core

    cases = [layered_volume(128, 128, nz, seed=s) for nz, s in ((9, 1), (12, 2), (7, 3), (9, 4))]
    cases += [layered_volume(64, 64, 10, seed=5, mask=False), layered_volume(64, 64, 6, seed=6)]
    its = 2
    alone = []
    for a in cases:
        s = _solver(a, precision)
        s.upload(a["field"])
        s.run_resident(its)
        alone.append(s.fetch(its))
        s.close()
    solvers = [_solver(a, precision) for a in cases]
    for s, a in zip(solvers, cases):
        s.upload(a["field"])
    _core.run_batch(solvers, its, max_lane=8)
    for s, a, ref in zip(solvers, cases, alone):
        got = s.fetch(its)
        for key in ("transmitted", "reflected"):
            assert np.array_equal(got[key], ref[key]), key
        for t_got, t_ref in zip(got["transmitted_iter"], ref["transmitted_iter"]):
            assert np.array_equal(t_got, t_ref)
        assert got["stored_values"] == ref["stored_values"]
    # and the batch is right: one case against the oracle
    a = cases[1]
    want = O.Solver(a["ny"], a["nx"], a["dy"], a["dx"], a["dz"], a["k0"], a["indices"], a["labels"],
                    a["index_before"], a["mask"]).run(a["field"], its)
    got = solvers[1].fetch(its)
    tol = 1e-10 if precision == "fp64" else 1e-5
    assert rel_l2(got["transmitted"], want["transmitted"]) < tol
    assert rel_l2(got["reflected"], want["reflected"]) < tol


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_interface_row_variants_give_the_same_bits(precision: str) -> None:
    """Interface steps run the row kernel's LDS-merge variant when a launch holds a boundary whose
    rows mostly select several label pairs, and the plain variant otherwise (merging through
    product plane 0 where a workgroup's rows mix pairs; rows of one pair take their product
    directly). Three 64² solves cross interfaces in the same launches: a thin cylinder (its disc
    rows, ~20 %, mix pairs), a plane tilted along x (every row mixes pairs, so the lane takes the
    merge variant) and one tilted along y (every row has one pair). Alone, the cylinder and the
    y-tilt run the plain variant; in the lane they run the merge variant. Same bits required."""
    import dataclasses

    from bwpm_amd import _core, workloads
    from bwpm_amd.specs import Above, Cylinder, Geometry

    base = workloads.tilted_3d(n=64, nz=24, iterations=2, geometry=True)
    dz, dx = base.dz, base.dx
    tan = float(np.tan(np.deg2rad(10.0)))
    probs = [dataclasses.replace(base, labels=Geometry(24, (Cylinder(1, 6 * dx, 10 * dz, 14 * dz),))),
             dataclasses.replace(base, labels=Geometry(24, (Above(1, 12 * dz, tan, 0.0),))),
             workloads.tilted_3d(n=64, nz=24, iterations=2, polarisation="y", geometry=True)]
    alone = []
    for p in probs:
        s = _core.GpuSolver(*p.solver_args(), precision=precision)
        try:
            alone.append(s.run(p.field, p.iterations))
        finally:
            s.close()
    solvers = [_core.GpuSolver(*p.solver_args(), precision=precision) for p in probs]
    try:
        for s, p in zip(solvers, probs):
            s.upload(p.field)
        _core.run_batch(solvers, 2)
        for i, (s, ref) in enumerate(zip(solvers, alone)):
            got = s.fetch(2)
            for key in ("transmitted", "reflected"):
                assert np.array_equal(got[key], ref[key]), (i, key)
    finally:
        for s in solvers:
            s.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_generic_kernels_narrow_and_wide_give_the_same_bits(precision: str) -> None:
    """Generic sizes (96² = 2^5·3 lean, 80×120 with a radix-5 stage): a solve alone launches few
    workgroups and takes the narrow geometry (less LDS, fewer rows per workgroup); in a lane of 28
    the launch is large enough for the wide one. Lines are transformed alike either way."""
    from bwpm_amd import _core

    for ny, nx in ((96, 96), (80, 120)):
        cases = [layered_volume(ny, nx, 8, seed=40 + i) for i in range(28)]
        alone = []
        for a in cases[:3]:
            s = _solver(a, precision)
            try:
                alone.append(s.run(a["field"], 2))
            finally:
                s.close()
        solvers = [_solver(a, precision) for a in cases]
        try:
            for s, a in zip(solvers, cases):
                s.upload(a["field"])
            _core.run_batch(solvers, 2, max_lane=28)
            for s, ref in zip(solvers[:3], alone):
                got = s.fetch(2)
                for key in ("transmitted", "reflected"):
                    assert np.array_equal(got[key], ref[key]), (ny, nx, key)
        finally:
            for s in solvers:
                s.close()


@pytest.mark.parametrize("shape", [(4096, 1), (2048, 2), (1, 4096), (512, 512)])
@pytest.mark.parametrize("precision", ["fp64", "mixed", "bfp", "fp32"])
def test_repeated_solves_are_bitwise_equal(shape: tuple[int, int], precision: str) -> None:
    """The same solve twice gives the same bits (the kernels read nothing they have not written and
    reduce in a fixed order), here at the line lengths of 4096 points whose FP64 kernels run with
    128 registers per thread."""
    ny, nx = shape
    a = layered_volume(ny, nx, 7, seed=ny + 3 * nx)
    s = _solver(a, precision)
    try:
        one = s.run(a["field"], 2)
        two = s.run(a["field"], 2)
    finally:
        s.close()
    for key in ("transmitted", "reflected"):
        assert np.array_equal(one[key], two[key]), key
