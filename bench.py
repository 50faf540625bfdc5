#!/usr/bin/env python3
"""Benchmark of the 3D vector BWPM slab-stepping path on one MI355X.

Workload (default): the tilted-interface 3D vector geometry of the reference's plain config
(λ = 1.55 µm, dx = dy = λ/6, dz = λ/4, interface 1.0 → 1.46 tilted 10°, pixel-centre labels,
Gaussian w₀ = 8 µm, 4 bidirectional iterations) at 512² × 1024 slabs, FP64 (complex128) bulk by
default. One step is one batch of `--batch` independent solves (E_x and E_y polarised alternately)
run on the GPU as one lane (every slab step is one column launch and one row launch for all solves
of the batch) with fields and state resident in HBM; the timed region covers `--steps` such batches.

Printed: one JSON line with grid-point·slab updates/s (whole job), the roofline of the slab step
(HIP-event durations of the column and the row kernel on the lane's stream; algorithmic bytes per
update = one read and one write of the (E_x, E_y) plane: 64 B in complex128, 32 B in complex64),
the plain one-solve-per-launch rate of the same solve on the same GPU, the accuracy of the run
(batch result against the solve alone; FP32 against the FP64 device path), the reference's other
configurations as single solves (2D, 3D 256² tilted, 3D 512² Born volume at reduced depth), the
served mix on the request queue (capacity, then p50/p99 latency of a Poisson stream at 70 % of it), and the
reference's CPU core timed on this host over a bounded sample of the same geometry.

    python bench.py [--steps K] [--warmup W] [--batch B] [--n 512] [--nz 1024] [--iterations 4]
                    [--precision fp64|mixed|fp32|fp32_plain]
"""

from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E
# algorithmic bytes per grid-point·slab update: read + write of the (E_x, E_y) plane per slab step
BYTES_PER_UPDATE = {"fp64": 64, "mixed": 32, "bfp": 32, "fp32": 32, "fp32_plain": 32}


def profiled_traffic(precision: str, n: int, batch: int) -> tuple[int | None, str | None]:
    """HBM bytes per launch pair from the committed rocprofv3 summary of the same configuration
    (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the copy-kernel calibration)."""
    tag = f"config: precision={precision} n={n} batch={batch}"
    for f in sorted((ROOT / "profiles").glob("*.md")):
        text = f.read_text()
        m = re.search(r"^traffic_bytes_per_launch_pair: (\d+)$", text, re.M)
        if m and tag in text:
            return int(m.group(1)), f.name
    return None, None


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / nb if nb > 0 else float(np.linalg.norm(a))


def cpu_baseline(n: int, nz_sample: int, threads: int) -> dict:
    """The reference's compiled CPU core (oracle/_ref) on a bounded sample of the workload, or the
    NumPy oracle when that build is absent."""
    from bwpm_amd import workloads

    prob = workloads.tilted_3d(n=n, nz=nz_sample, iterations=1)
    args = (prob.ny, prob.nx, prob.dy, prob.dx, prob.dz, prob.k0, np.asarray(prob.indices, complex),
            prob.labels.astype(np.int32), prob.index_before, None)
    ref_dir = ROOT / "oracle" / "_ref"
    so = sorted(ref_dir.glob("_core*.so")) if ref_dir.exists() else []
    if so:
        import importlib.util

        spec = importlib.util.spec_from_file_location("_core", so[0])
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        solver = mod.Solver(*args, threads)
        kind, cores = "reference", threads
    else:
        from oracle import bwpm_oracle

        solver = bwpm_oracle.Solver(*args)
        kind, cores = "port", 1
    t0 = time.perf_counter()
    solver.run(prob.field, prob.iterations, [])
    dt = time.perf_counter() - t0
    return {"value": prob.updates / dt, "unit": "grid-point·slab updates/s", "cores": cores, "kind": kind,
            "sample": f"tilted_3d {n}x{n}x{nz_sample} slabs, 1 iteration ({prob.updates} updates) in {dt:.2f} s"}


def run_solve(p, precision: str) -> tuple[dict, float]:
    """One solve resident on the device: result and the updates/s of its timed run (after a warm run)."""
    from bwpm_amd import _core

    s = _core.GpuSolver(*p.solver_args(), fluctuation=p.fluctuation, precision=precision)
    try:
        s.upload(p.field)
        s.run_resident(p.iterations)
        t0 = time.perf_counter()
        s.run_resident(p.iterations)
        dt = time.perf_counter() - t0
        return s.fetch(p.iterations), p.updates / dt
    finally:
        s.close()


def precision_mode(probs: list, precision: str, iterations: int, ref: dict) -> dict:
    """The headline lane in another precision: updates/s over one timed batch after one warm-up,
    the kernel durations, and the relative L2 error of solve 0 against the FP64 device result."""
    from bwpm_amd import _core

    solvers = []
    try:
        for p in probs:
            s = _core.GpuSolver(*p.solver_args(), precision=precision)
            s.upload(p.field)
            solvers.append(s)
        _core.run_batch(solvers, iterations, max_lane=len(solvers))
        t0 = time.perf_counter()
        _core.run_batch(solvers, iterations, max_lane=len(solvers))
        rate = sum(p.updates for p in probs) / (time.perf_counter() - t0)
        kt = _core.run_batch(solvers, iterations, max_lane=len(solvers), time_kernels=True)
        got = solvers[0].fetch(iterations)
    finally:
        for s in solvers:
            s.close()
    return {"updates_per_s": rate,
            "k_col_us": round(1e3 * kt["col_ms"] / max(kt["col_launches"], 1), 3),
            "k_row_us": round(1e3 * kt["row_ms"] / max(kt["row_launches"], 1), 3),
            "rel_l2_transmitted": rel_l2(got["transmitted"], ref["transmitted"]),
            "rel_l2_reflected": rel_l2(got["reflected"], ref["reflected"])}


def small_lane(count: int = 64, n: int = 128, nz: int = 256, iterations: int = 2) -> dict:
    """Batched over plain on small grids: `count` tilted-interface solves of n² × nz in one lane
    against the same solves one at a time (one solve per launch, synchronised per solve)."""
    from bwpm_amd import _core, workloads

    probs = [workloads.tilted_3d(n=n, nz=nz, iterations=iterations, polarisation="xy"[i % 2]) for i in range(count)]
    solvers = []
    try:
        for p in probs:
            s = _core.GpuSolver(*p.solver_args())
            s.upload(p.field)
            solvers.append(s)
        _core.run_batch(solvers, iterations, max_lane=count)
        t0 = time.perf_counter()
        _core.run_batch(solvers, iterations, max_lane=count)
        batched = sum(p.updates for p in probs) / (time.perf_counter() - t0)
        solvers[0].run_resident(iterations)
        t0 = time.perf_counter()
        for s in solvers:
            s.run_resident(iterations)
        plain = sum(p.updates for p in probs) / (time.perf_counter() - t0)
    finally:
        for s in solvers:
            s.close()
    return {"workload": f"{count} tilted-interface solves {n}x{n}x{nz}, {iterations} iterations, FP64",
            "batched_updates_per_s": batched, "plain_updates_per_s": plain, "batched_over_plain": batched / plain}


def uneven_batch(count: int) -> dict:
    """Config 4: `count` uneven Born solves (128² … 512², 100 … 1000 slabs, per-solve λ, angle, σ and
    seed; δn generated on the device) submitted at once to the request queue, complex64."""
    from bwpm_amd import serve, workloads

    probs = workloads.uneven_batch(count)
    with serve.Server(slots=16, chunk=64, precision="fp32") as srv:
        t0 = time.perf_counter()
        tickets = [srv.submit(p) for p in probs]
        for t in tickets:
            srv.result(t)
        span = time.perf_counter() - t0
        st = srv.stats()
    return {"solves": count, "seconds": round(span, 2), "updates_per_s": sum(p.updates for p in probs) / span,
            "p50_ms": round(st["p50_ms"], 1), "p99_ms": round(st["p99_ms"], 1), "failed": int(st["failed"]),
            "peak_hbm_gb": round(st["peak_bytes"] / 1e9, 2)}


def configs(c3_nz: int, c4_count: int, which: str = "c1,c2,c3,c4") -> dict:
    """The reference's other configurations: config 1 (2D, 1024 points, 200 slabs, 3 iterations),
    config 2 (3D 256² × 512 tilted interface, 4 iterations, E_x and E_y) and config 3 (3D 512² Born
    volume, `c3_nz` slabs, 6 iterations) as single solves; config 4 (uneven Born batch) on the
    queue. Errors of the mixed and FP32 modes are relative L2 against the FP64 device result
    (itself within 1e-10 of the oracle). `which` selects configurations (comma-separated)."""
    from bwpm_amd import workloads

    sel = {w.strip() for w in which.split(",")}
    out: dict = {}
    if "c1" in sel:
        p1 = workloads.plain_2d(ny=1024, nz=200, iterations=3)
        _, r1 = run_solve(p1, "fp64")
        out["c1_2d_1024x200_it3"] = {"fp64_updates_per_s": r1}

    def against(p, ref: dict) -> dict:
        # the same solve with complex64 streaming planes (FP64 arithmetic) and in FP32 arithmetic
        out_m = {}
        for m in ("mixed", "fp32"):
            r, rate = run_solve(p, m)
            out_m[m] = {"updates_per_s": rate, "rel_l2_transmitted": rel_l2(r["transmitted"], ref["transmitted"]),
                        "rel_l2_reflected": rel_l2(r["reflected"], ref["reflected"])}
        return out_m

    if "c2" in sel:
        c2 = {}
        for pol in ("x", "y"):
            p2 = workloads.tilted_3d(n=256, nz=512, iterations=4, polarisation=pol)
            a64, r64 = run_solve(p2, "fp64")
            c2[f"E{pol}"] = {"fp64_updates_per_s": r64, **against(p2, a64)}
        out["c2_tilted_256x256x512_it4"] = c2
    if "c3" in sel:
        # config 3 at full depth: the δn volume is generated on the device (a (1024, 512, 512)
        # complex array would be 4 GiB on the host)
        p3 = workloads.born_3d_synth(n=512, nz=c3_nz, iterations=6)
        b64, q64 = run_solve(p3, "fp64")
        out[f"c3_born_512x512x{c3_nz}_it6"] = {
            "fp64_updates_per_s": q64, "fp64_gmres_iterations": b64["gmres_iterations"], **against(p3, b64),
            "input": "Gaussian-correlated δn (σ = 0.01, correlation 1 µm, seed 3) generated on the device"}
    if c4_count > 0 and "c4" in sel:
        out[f"c4_uneven_born_{c4_count}"] = uneven_batch(c4_count)
    return out


def served(count: int, precision: str) -> dict:
    """The served mix on the request queue: a burst of `count` requests measures capacity, then a
    Poisson stream of `count` other requests at 70 % of it gives latency under load."""
    from bwpm_amd import serve

    burst = [serve.served_request(1000 + i) for i in range(count)]
    with serve.Server(slots=16, chunk=64, precision=precision) as srv:
        t0 = time.perf_counter()
        tickets = [srv.submit(p) for p in burst]
        for t in tickets:
            srv.result(t)
        span = time.perf_counter() - t0
        cap = srv.stats()
    rate = 0.7 * count / span
    with serve.Server(slots=16, chunk=64, precision=precision) as srv:
        rep = serve.run_stream(srv, (serve.served_request(2000 + i) for i in range(count)), rate_hz=rate, seed=1)
    st = rep.stats
    return {
        "workload": "served mix: 2D 512-4096 points and 3D vector 128^2-512^2, 200-1024 slabs, planar / tilted / "
                    "silver-layer / index-fluctuation geometries, 1-3 iterations (seeded)",
        "requests": count,
        "capacity_req_per_s": round(count / span, 3),
        "capacity_updates_per_s": float(sum(p.updates for p in burst) / span),
        "burst_p50_ms": round(cap["p50_ms"], 1),
        "burst_p99_ms": round(cap["p99_ms"], 1),
        "offered_rate_hz": round(rate, 3),
        "p50_ms": round(st["p50_ms"], 1),
        "p99_ms": round(st["p99_ms"], 1),
        "mean_ms": round(st["mean_ms"], 1),
        "updates_per_s": st["updates_per_s"],
        "completed": int(st["completed"]),
        "failed": int(st["failed"]) + len(rep.failures),
        "peak_hbm_gb": round(st["peak_bytes"] / 1e9, 2),
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--nz", type=int, default=1024)
    ap.add_argument("--iterations", type=int, default=4)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed", "bfp", "fp32", "fp32_plain"])
    ap.add_argument("--absorber", type=float, default=0.0,
                    help="absorbing edge zone of the headline geometry in wavelengths (0: none)")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    ap.add_argument("--cpu-nz", type=int, default=48)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-checks", action="store_true", help="skip the plain-path and accuracy runs")
    ap.add_argument("--no-modes", action="store_true", help="skip the mixed and FP32 lanes of the headline")
    ap.add_argument("--serve-requests", type=int, default=24,
                    help="requests of the served mix (0: no serving run)")
    ap.add_argument("--c3-nz", type=int, default=1024, help="slabs of the Born configuration (0: skip configs)")
    ap.add_argument("--c4-count", type=int, default=64, help="solves of the uneven Born batch (0: skip)")
    ap.add_argument("--configs", default="c1,c2,c3,c4", help="configurations to run (comma-separated)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU figures are not produced",
                          "n_gpus": args.gpus}))
        return 2

    from bwpm_amd import _core, workloads

    probs = [workloads.tilted_3d(n=args.n, nz=args.nz, iterations=args.iterations,
                                 polarisation="x" if i % 2 == 0 else "y", absorber_wavelengths=args.absorber)
             for i in range(args.batch)]
    solvers = []
    for p in probs:
        s = _core.GpuSolver(*p.solver_args(), precision=args.precision)
        s.upload(p.field)
        solvers.append(s)

    def one_step() -> None:
        # one lane: every slab step is one column and one row launch for all solves of the batch
        _core.run_batch(solvers, args.iterations, max_lane=args.batch)

    for _ in range(args.warmup):
        one_step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one_step()
    elapsed = time.perf_counter() - t0
    updates = sum(p.updates for p in probs) * args.steps
    value = updates / elapsed
    ms_per_step = 1e3 * elapsed / max(args.steps, 1)

    # kernel durations inside the lane: HIP events around every launch on the lane's stream
    kt = _core.run_batch(solvers, args.iterations, max_lane=args.batch, time_kernels=True)
    bpu = BYTES_PER_UPDATE[args.precision]
    npix = args.n * args.n * args.batch  # pixels one launch processes
    col_us = 1e3 * kt["col_ms"] / max(kt["col_launches"], 1)
    row_us = 1e3 * kt["row_ms"] / max(kt["row_launches"], 1)
    step_us = col_us + row_us
    achieved = bpu * npix / (step_us * 1e-6) / 1e9
    traffic, traffic_src = profiled_traffic(args.precision, args.n, args.batch)
    copy_bytes = 1 << 30
    copy_ms = _core.copy_probe(copy_bytes, 10)
    roofline = {
        "bound": "hbm",
        "kernel": "k_col_fast + k_row_fast (one slab step of the lane)",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBPS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBPS, 5),
        "traffic": traffic,
        "traffic_source": traffic_src,
        "algorithmic_bytes_per_launch_pair": bpu * npix,
        "k_col_us": round(col_us, 3),
        "k_row_us": round(row_us, 3),
        "launch_pairs": kt["col_launches"],
        "copy_probe_GBps": round(2 * copy_bytes / (copy_ms * 1e-3) / 1e9, 1),
    }

    checks = None
    if not args.no_checks:
        batch_res = solvers[0].fetch(args.iterations)
        # the plain path: the same solve alone, one solve per launch, synchronised per solve
        solvers[0].run_resident(args.iterations)
        t0 = time.perf_counter()
        solvers[0].run_resident(args.iterations)
        plain = probs[0].updates / (time.perf_counter() - t0)
        alone = solvers[0].fetch(args.iterations)
        # the same solve through host buffers (field copied in, results copied out over PCIe)
        t0 = time.perf_counter()
        solvers[0].run(probs[0].field, args.iterations)
        host_rate = probs[0].updates / (time.perf_counter() - t0)
        solvers[0].upload(probs[0].field)
        checks = {
            "plain_one_solve_per_launch": plain,
            "one_solve_with_host_buffers": host_rate,
            "batched_over_plain": value / plain,
            "batch_vs_alone_bitwise": bool(np.array_equal(batch_res["transmitted"], alone["transmitted"])
                                           and np.array_equal(batch_res["reflected"], alone["reflected"])),
        }
        if args.precision != "fp64":
            ref = _core.GpuSolver(*probs[0].solver_args(), precision="fp64")
            ref.upload(probs[0].field)
            ref.run_resident(args.iterations)
            r64 = ref.fetch(args.iterations)
            ref.close()
            checks["rel_l2_vs_fp64_device"] = {"transmitted": rel_l2(alone["transmitted"], r64["transmitted"]),
                                               "reflected": rel_l2(alone["reflected"], r64["reflected"])}

    for s in solvers:
        s.close()
    solvers = []
    if checks is not None:
        checks["small_grids"] = small_lane()
    modes = None
    if checks is not None and args.precision == "fp64" and not args.no_modes:
        # the same lane with complex64 streaming planes (FP64 arithmetic) and in FP32 arithmetic
        modes = {m: precision_mode(probs, m, args.iterations, alone) for m in ("mixed", "fp32")}
    other = configs(args.c3_nz, args.c4_count, args.configs) if args.c3_nz > 0 else None
    serving = None
    if args.serve_requests > 0:
        serving = served(args.serve_requests, args.precision)

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads if args.cpu_threads > 0 else min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(args.n, args.cpu_nz, threads)
    line = {
        "metric": "grid-point·slab updates/s (3D vector 512²)",
        "value": value,
        "unit": "grid-point·slab updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "c128", "mixed": "c64 planes, f64 arithmetic"}.get(args.precision, "c64"),
        "data": "synthetic",
        "config": {
            "workload": f"tilted-interface 3D vector BWPM {args.n}x{args.n}x{args.nz} slabs, "
                        f"{args.iterations} iterations, batch of {args.batch} solves in one lane",
            "n": args.n, "nz": args.nz, "iterations": args.iterations, "batch": args.batch,
            "precision": args.precision, "updates_per_solve": probs[0].updates,
        },
        "roofline": roofline,
        "checks": checks,
        "precision_modes": modes,
        "configs": other,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    for s in solvers:
        s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
