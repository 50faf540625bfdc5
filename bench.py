#!/usr/bin/env python3
"""Benchmark of the 3D vector BWPM slab-stepping path on one MI355X.

Workload (default): the tilted-interface 3D vector geometry of the reference's plain config
(λ = 1.55 µm, dx = dy = λ/6, dz = λ/4, interface 1.0 → 1.46 tilted 10°, pixel-centre labels,
Gaussian w₀ = 8 µm, 4 bidirectional iterations) at 512² × 1024 slabs with the reference's absorbing
edge zone of 20 wavelengths (its validity rule; without it the periodic 512² window lets refracted
waves circulate and the fields grow about 400× per iteration, in the reference's own core as here).
Precision `mixed` by default: the two streaming planes of every slab step stored as complex64 in
HBM, FP64 arithmetic and operator tables. One step is one batch of `--batch` independent solves
(E_x and E_y polarised alternately) run on the GPU as one lane (every slab step is one column launch
and one row launch for all solves of the batch) with fields and state resident in HBM; the timed
region covers `--steps` such batches.

Printed: one JSON line with grid-point·slab updates/s (whole job), the roofline of the slab step
(HIP-event durations of the column and the row kernel on the lane's stream; algorithmic bytes per
update = one read and one write of the (E_x, E_y) plane: 64 B in complex128, 32 B in complex64 or
the block-floating-point format), and for every result the relative field L2 error, reflectance R,
transmittance T and A = 1 − R − T against the FP64 device result of the same solve (the FP64
device path equals the FP64 oracle and the reference's C++ core to 1e-10 in the GPU tests), with
|ΔR| and |ΔT|. Also: the plain one-solve-per-launch rates, the other precision modes, the
reference's other configurations (2D; 3D 256² tilted; 3D 512² Born volume; the uneven Born batch),
the served mix on the request queue, and the reference's compiled CPU core timed on this host (all
cores and one core, one warm-up run and the median of five) on bounded samples.

    python bench.py [--steps K] [--warmup W] [--batch B] [--n 512] [--nz 1024] [--iterations 4]
                    [--precision mixed|bfp|fp32|fp32_plain|fp64] [--absorber 20]
"""

from __future__ import annotations

import argparse
import json
import os
import re
import statistics
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E
# algorithmic bytes per grid-point·slab update: read + write of the (E_x, E_y) plane per slab step
BYTES_PER_UPDATE = {"fp64": 64, "mixed": 32, "bfp": 32, "fp32": 32, "fp32_plain": 32}
# arithmetic type of each mode, and the storage of its streaming planes
ARITHMETIC = {"fp64": "f64", "mixed": "f64", "bfp": "f64", "fp32": "f32", "fp32_plain": "f32"}
PLANES = {"fp64": "complex128", "mixed": "complex64", "bfp": "block-floating-point (8 B per complex value)",
          "fp32": "complex64", "fp32_plain": "complex64"}
NARROW = ("mixed", "bfp", "fp32")


def profiled_traffic(precision: str, n: int, batch: int, absorber: float) -> tuple[int | None, str | None]:
    """HBM bytes per launch pair from the committed rocprofv3 summary of the same configuration
    (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the copy-kernel calibration)."""
    tag = f"config: precision={precision} n={n} batch={batch} absorber={absorber:g}"
    for f in sorted((ROOT / "profiles").glob("*.md")):
        text = f.read_text()
        m = re.search(r"^traffic_bytes_per_launch_pair: (\d+)$", text, re.M)
        if m and tag in text:
            return int(m.group(1)), f.name
    return None, None


_T0 = time.perf_counter()


def progress(msg: str) -> None:
    """A progress line on stderr (stdout carries only the JSON line)."""
    print(f"bench [{time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / nb if nb > 0 else float(np.linalg.norm(a))


# ---- reflectance, transmittance and errors ---------------------------------------------------

def exit_index(p) -> complex:
    """Index of the medium after the volume: the label of the last slab at the centre pixel
    (src/bwpm/adapter.py:267-284)."""
    from bwpm_amd import _core
    from bwpm_amd.specs import Geometry, Uniform

    lab = p.labels
    if isinstance(lab, Uniform):
        label = lab.label
    elif isinstance(lab, Geometry):
        last = Geometry(1, lab.prims, lab.background, lab.z0 + (lab.nz - 1) * p.dz)
        label = int(_core.geometry_labels(last, p.ny, p.nx, p.dy, p.dx, p.dz, len(p.indices))[0, p.ny // 2, p.nx // 2])
    else:
        label = int(np.asarray(lab)[-1, p.ny // 2, p.nx // 2])
    return complex(p.indices[label])


def r_t_a(p, res: dict) -> dict:
    """R and T as Poynting fluxes over the incident flux (device FP64 reductions), A = 1 − R − T
    (the energy residual of src/bwpm/fast.py:146-155)."""
    from bwpm_amd import _core

    def flux(f, n, d):
        return _core.power_flux(f, p.dy, p.dx, p.k0, n, d)

    p_in = flux(p.field, p.index_before, 1)
    r = flux(res["reflected"], p.index_before, -1) / p_in
    t = flux(res["transmitted"], exit_index(p), 1) / p_in
    return {"R": r, "T": t, "A": 1.0 - r - t}


def accuracy(p, got: dict, ref: dict, ref_rta: dict | None = None) -> dict:
    """Field L2 errors and R/T/A of `got` against the FP64 result `ref` of the same solve."""
    g = r_t_a(p, got)
    r = ref_rta or r_t_a(p, ref)
    return {"rel_l2_transmitted": rel_l2(got["transmitted"], ref["transmitted"]),
            "rel_l2_reflected": rel_l2(got["reflected"], ref["reflected"]),
            "R": g["R"], "T": g["T"], "A": g["A"],
            "abs_dR": abs(g["R"] - r["R"]), "abs_dT": abs(g["T"] - r["T"]), "abs_dA": abs(g["A"] - r["A"])}


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


def iteration(res: dict, i: int) -> dict:
    return {"transmitted": res["transmitted_iter"][i], "reflected": res["reflected_iter"][i]}


def with_modes(p, modes=NARROW) -> dict:
    """The solve in FP64 (rate, R, T, A, and R and T after each iteration) and in each narrow mode
    (rate and accuracy against FP64, also after each iteration: an iteration that does not settle
    amplifies the rounding of the narrow modes with its own growth)."""
    ref, rate = run_solve(p, "fp64")
    ref_rta = r_t_a(p, ref)
    its = len(ref.get("reflected_iter", []))
    per_it = [r_t_a(p, iteration(ref, i)) for i in range(its)]
    out = {"fp64": {"updates_per_s": rate, **ref_rta, "R_per_iteration": [e["R"] for e in per_it],
                    "T_per_iteration": [e["T"] for e in per_it]}}
    if "gmres_iterations" in ref:
        out["fp64"]["gmres_iterations"] = ref["gmres_iterations"]
    for m in modes:
        got, r = run_solve(p, m)
        out[m] = {"updates_per_s": r, **accuracy(p, got, ref, ref_rta)}
        acc_it = [accuracy(p, iteration(got, i), iteration(ref, i), per_it[i]) for i in range(its)]
        for key in ("rel_l2_transmitted", "rel_l2_reflected", "abs_dR", "abs_dT"):
            out[m][key + "_per_iteration"] = [e[key] for e in acc_it]
        if "gmres_iterations" in got:
            out[m]["gmres_iterations"] = got["gmres_iterations"]
    return out


# ---- the headline lane and its variants --------------------------------------------------------

def lane(probs: list, precision: str, iterations: int, warmup: int, steps: int) -> tuple[list, float, dict, list]:
    """Solvers of `probs` in one lane: the seconds of `steps` timed batches after `warmup`, the
    HIP-event kernel durations of one more batch, and the results of solves 0 and 1. The caller
    closes the returned solvers."""
    from bwpm_amd import _core

    solvers = []
    try:
        for p in probs:
            s = _core.GpuSolver(*p.solver_args(), precision=precision)
            s.upload(p.field)
            solvers.append(s)
        for _ in range(warmup):
            _core.run_batch(solvers, iterations, max_lane=len(solvers))
        t0 = time.perf_counter()
        for _ in range(steps):
            _core.run_batch(solvers, iterations, max_lane=len(solvers))
        elapsed = time.perf_counter() - t0
        kt = _core.run_batch(solvers, iterations, max_lane=len(solvers), time_kernels=True)
        results = [solvers[i].fetch(iterations) for i in range(min(2, len(solvers)))]
    except BaseException:
        for s in solvers:
            s.close()
        raise
    return solvers, elapsed, kt, results


def kernel_us(kt: dict) -> tuple[float, float]:
    return (1e3 * kt["col_ms"] / max(kt["col_launches"], 1), 1e3 * kt["row_ms"] / max(kt["row_launches"], 1))


def small_lane(count: int = 64, n: int = 128, nz: int = 256, iterations: int = 2) -> dict:
    """Batched over plain on small grids: `count` tilted-interface solves of n² × nz in one lane
    against the same solves one at a time (one solve per launch, synchronised per solve), FP64."""
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
        in_batch = solvers[0].fetch(iterations)
        solvers[0].run_resident(iterations)
        t0 = time.perf_counter()
        for s in solvers:
            s.run_resident(iterations)
        plain = sum(p.updates for p in probs) / (time.perf_counter() - t0)
        alone = solvers[0].fetch(iterations)
    finally:
        for s in solvers:
            s.close()
    same = bool(np.array_equal(in_batch["transmitted"], alone["transmitted"])
                and np.array_equal(in_batch["reflected"], alone["reflected"]))
    return {"workload": f"{count} tilted-interface solves {n}x{n}x{nz}, {iterations} iterations, FP64",
            "batched_updates_per_s": batched, "plain_updates_per_s": plain, "batched_over_plain": batched / plain,
            "batch_vs_alone_bitwise": same}


# ---- the reference's other configurations ------------------------------------------------------

def uneven_batch(count: int, precision: str, check: int = 4) -> dict:
    """Config 4: `count` uneven Born solves (128² … 512², 100 … 1000 slabs, per-solve λ, angle, σ and
    seed; δn generated on the device) submitted at once to the request queue. Accuracy: the
    `check` smallest solves again alone in FP64 against their results from the queue."""
    from bwpm_amd import serve, workloads

    probs = workloads.uneven_batch(count)
    order = sorted(range(count), key=lambda i: probs[i].updates)[:check]
    with serve.Server(slots=16, chunk=64, precision=precision) as srv:
        t0 = time.perf_counter()
        tickets = [srv.submit(p) for p in probs]
        kept = {}
        for i, t in enumerate(tickets):
            r = srv.result(t)
            if i in order:
                kept[i] = r
        span = time.perf_counter() - t0
        st = srv.stats()
    acc = {}
    for i in order:
        ref, _ = run_solve(probs[i], "fp64")
        acc[probs[i].name] = accuracy(probs[i], kept[i], ref)
    return {"solves": count, "precision": precision, "seconds": round(span, 2),
            "updates_per_s": sum(p.updates for p in probs) / span,
            "p50_ms": round(st["p50_ms"], 1), "p99_ms": round(st["p99_ms"], 1), "failed": int(st["failed"]),
            "peak_hbm_gb": round(st["peak_bytes"] / 1e9, 2),
            "accuracy_vs_fp64_alone": acc}


def configs(c3_nz: int, c4_count: int, precision: str, which: str = "c1,c2,c3,c4") -> dict:
    """Config 1 (2D, 1024 points, 200 slabs, 3 iterations), config 2 (3D 256² × 512 tilted
    interface, 4 iterations, E_x and E_y) and config 3 (3D 512² Born volume, `c3_nz` slabs, 6
    iterations) as single solves in FP64 and in each narrow mode; config 4 (uneven Born batch) on
    the queue in `precision` (bfp by default: the reflected fields of Born volumes are 10⁻³ … 10⁻²
    of the forward field and partly cancel, which lifts complex64's rounding to about 1e-5 of them,
    while bfp keeps them near 2e-6 at the same 8 bytes). `which` selects configurations."""
    from bwpm_amd import workloads

    sel = {w.strip() for w in which.split(",")}
    out: dict = {}
    if "c1" in sel:
        out["c1_2d_1024x200_it3"] = with_modes(workloads.plain_2d(ny=1024, nz=200, iterations=3))
        progress("config 1 done")
    if "c2" in sel:
        out["c2_tilted_256x256x512_it4"] = {
            f"E{pol}": with_modes(workloads.tilted_3d(n=256, nz=512, iterations=4, polarisation=pol))
            for pol in ("x", "y")}
        progress("config 2 done")
    if "c3" in sel and c3_nz > 0:
        # the δn volume is generated on the device (a (1024, 512, 512) complex array would be
        # 4 GiB on the host); the transverse edges carry the reference's absorbing zone of 20
        # wavelengths (its validity rule): in the bare periodic window the iteration diverges
        # from the fourth iteration on, in FP64 as in every narrow mode
        out[f"c3_born_512x512x{c3_nz}_it6"] = {
            **with_modes(workloads.born_3d_synth(n=512, nz=c3_nz, iterations=6, absorber_wavelengths=20.0)),
            "input": "Gaussian-correlated δn (σ = 0.01, correlation 1 µm, seed 3) generated on the device; "
                     "20-wavelength absorbing zone"}
        progress("config 3 done")
    if c4_count > 0 and "c4" in sel:
        out[f"c4_uneven_born_{c4_count}"] = uneven_batch(c4_count, precision)
        progress("config 4 done")
    return out


def served(count: int, precision: str, check: int = 3) -> dict:
    """The served mix on the request queue: a burst of `count` requests measures capacity, then a
    Poisson stream of `count` other requests at 70 % of it gives latency under load. Accuracy: the
    first `check` requests of the burst again alone in FP64."""
    from bwpm_amd import serve

    burst = [serve.served_request(1000 + i) for i in range(count)]
    with serve.Server(slots=16, chunk=64, precision=precision) as srv:
        t0 = time.perf_counter()
        tickets = [srv.submit(p) for p in burst]
        kept = {}
        for i, t in enumerate(tickets):
            r = srv.result(t)
            if i < check:
                kept[i] = r
        span = time.perf_counter() - t0
        cap = srv.stats()
    rate = 0.7 * count / span
    with serve.Server(slots=16, chunk=64, precision=precision) as srv:
        rep = serve.run_stream(srv, (serve.served_request(2000 + i) for i in range(count)), rate_hz=rate, seed=1)
    st = rep.stats
    acc = {}
    for i in range(min(check, count)):
        ref, _ = run_solve(burst[i], "fp64")
        acc[burst[i].name] = accuracy(burst[i], kept[i], ref)
    return {
        "workload": "served mix: 2D 512-4096 points and 3D vector 128^2-512^2, 200-1024 slabs, planar / tilted / "
                    "silver-layer / index-fluctuation geometries, 1-3 iterations (seeded)",
        "precision": precision,
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
        "accuracy_vs_fp64_alone": acc,
    }


# ---- CPU baseline ------------------------------------------------------------------------------

def fft_plans_chosen() -> list:
    """The radix plans the plan tuner chose in this process (per precision, direction and line
    length: 0 default, 1 alternative, with the timed launch of each in µs)."""
    from bwpm_amd import _core

    return _core.fft_plans()


def host_cores() -> dict:
    """Physical cores from lscpu, its CPU model, and the CPUs this process may run on."""
    info: dict = {"allowed_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        pairs = {tuple(line.split(",")) for line in out.splitlines() if line and not line.startswith("#")}
        info["physical_cores"] = len(pairs)
        model = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        m = re.search(r"^Model name:\s*(.+)$", model, re.M)
        if m:
            info["model"] = m.group(1).strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(reps: int = 5) -> dict:
    """The reference's compiled C++ CPU core (oracle/_ref, built from its own sources) on bounded
    samples of the headline and of configs 1–3, on all CPUs this process may use and on one core:
    one warm-up run (which also fills its operator cache), then the median of `reps` timed runs."""
    import importlib.util

    from bwpm_amd import workloads

    so = sorted((ROOT / "oracle" / "_ref").glob("_core*.so")) if (ROOT / "oracle" / "_ref").exists() else []
    if not so:
        return {"error": "oracle/_ref not built (oracle/build_ref.sh)"}
    spec = importlib.util.spec_from_file_location("_core", so[0])
    ref = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ref)
    cores = host_cores()
    env_threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    all_threads = min(cores["allowed_cpus"], env_threads) if env_threads > 0 else cores["allowed_cpus"]
    samples = {
        "headline": (workloads.tilted_3d(n=512, nz=2, iterations=1, absorber_wavelengths=20.0),
                     "tilted 512x512, 2 slabs, 1 iteration, 20-wavelength absorber"),
        "c1": (workloads.plain_2d(ny=1024, nz=200, iterations=3), "2D 1024 points, 200 slabs, 3 iterations (full)"),
        "c2": (workloads.tilted_3d(n=256, nz=8, iterations=1), "tilted 256x256, 8 slabs, 1 iteration"),
        "c3": (workloads.born_3d(n=512, nz=1, iterations=1), "Born 512x512, 1 slab, 1 iteration"),
    }

    def timed(p, threads: int) -> tuple[float, list[float]]:
        s = ref.Solver(p.ny, p.nx, p.dy, p.dx, p.dz, p.k0, np.asarray(p.indices, complex),
                       np.asarray(p.labels, dtype=np.int32), complex(p.index_before), p.mask, threads,
                       fluctuation=p.fluctuation)
        ts = []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            s.run(p.field, p.iterations, [])
            ts.append(time.perf_counter() - t0)
        return p.updates / statistics.median(ts[1:]), ts

    out: dict = {}
    for key, (p, what) in samples.items():
        entry = {"sample": what, "updates": p.updates}
        for label, threads in (("all_cores", all_threads), ("one_core", 1)):
            rate, ts = timed(p, threads)
            entry[label] = {"threads": threads, "updates_per_s": rate, "seconds": [round(t, 4) for t in ts]}
        out[key] = entry
    head = out["headline"]["all_cores"]
    return {"value": head["updates_per_s"], "unit": "grid-point·slab updates/s", "cores": all_threads,
            "kind": "reference",
            "sample": f"{samples['headline'][1]} ({samples['headline'][0].updates} updates); one warm-up run, "
                      f"median of {reps}",
            "host": cores, "configs": out}


# ---- main --------------------------------------------------------------------------------------

def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=24)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--nz", type=int, default=1024)
    ap.add_argument("--iterations", type=int, default=4)
    ap.add_argument("--precision", default="mixed", choices=list(BYTES_PER_UPDATE))
    ap.add_argument("--absorber", type=float, default=20.0,
                    help="absorbing edge zone of the headline geometry in wavelengths (0: none)")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-checks", action="store_true", help="skip the accuracy and plain-path runs")
    ap.add_argument("--no-modes", action="store_true", help="skip the lanes of the other precision modes")
    ap.add_argument("--serve-requests", type=int, default=24,
                    help="requests of the served mix (0: no serving run)")
    ap.add_argument("--c3-nz", type=int, default=1024, help="slabs of the Born configuration (0: skip it)")
    ap.add_argument("--c4-count", type=int, default=64, help="solves of the uneven Born batch (0: skip)")
    ap.add_argument("--c4-precision", default="bfp", choices=list(BYTES_PER_UPDATE),
                    help="precision mode of the uneven Born batch")
    ap.add_argument("--configs", default="c1,c2,c3,c4", help="configurations to run (comma-separated; '' for none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU figures are not produced",
                          "n_gpus": args.gpus}))
        return 2

    from bwpm_amd import _core, workloads

    # E_x and E_y polarised solves alternate; solves of one polarisation share one problem (the
    # solvers copy their inputs)
    pol = {c: workloads.tilted_3d(n=args.n, nz=args.nz, iterations=args.iterations, polarisation=c,
                                  absorber_wavelengths=args.absorber) for c in "xy"}
    probs = [pol["x" if i % 2 == 0 else "y"] for i in range(args.batch)]
    solvers, elapsed, kt, batch_res = lane(probs, args.precision, args.iterations, args.warmup, args.steps)
    progress(f"headline lane timed ({elapsed:.2f} s for {args.steps} steps)")
    updates = sum(p.updates for p in probs) * args.steps
    value = updates / elapsed
    ms_per_step = 1e3 * elapsed / max(args.steps, 1)

    # kernel durations inside the lane: HIP events around every launch on the lane's stream
    bpu = BYTES_PER_UPDATE[args.precision]
    npix = args.n * args.n * args.batch  # pixels one launch processes
    col_us, row_us = kernel_us(kt)
    achieved = bpu * npix / ((col_us + row_us) * 1e-6) / 1e9
    traffic, traffic_src = profiled_traffic(args.precision, args.n, args.batch, args.absorber)
    copy_bytes = 1 << 30
    copy_ms = _core.copy_probe(copy_bytes, 10)
    roofline = {
        "bound": "hbm",
        "kernel": "slab step: column + row kernel (k_col_plain/k_row_plain on plain steps, k_col_fast/k_row_fast "
                  "at interfaces and pass ends), averaged over every launch of one batch",
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
    headline_acc = None
    modes = None
    try:
        if not args.no_checks:
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
                "batch_vs_alone_bitwise": bool(np.array_equal(batch_res[0]["transmitted"], alone["transmitted"])
                                               and np.array_equal(batch_res[0]["reflected"], alone["reflected"])),
            }
    finally:
        for s in solvers:
            s.close()
        solvers = []
    if checks is not None:
        # accuracy of the batch's E_x and E_y solves against the FP64 device result of each
        refs = {c: run_solve(pol[c], "fp64") for c in "xy"}
        ref_rta = {c: r_t_a(pol[c], refs[c][0]) for c in "xy"}
        headline_acc = {f"E{c}": accuracy(pol[c], batch_res[i], refs[c][0], ref_rta[c]) for i, c in enumerate("xy")}
        headline_acc["fp64_reference"] = {f"E{c}": ref_rta[c] for c in "xy"}
        # the plain one-solve-per-launch FP32 path (unbatched complex64, synchronised per solve)
        _, plain32 = run_solve(pol["x"], "fp32")
        checks["plain_fp32_one_solve_per_launch"] = plain32
        checks["batched_over_plain_fp32"] = value / plain32
        checks["plain_fp64_one_solve_per_launch"] = refs["x"][1]
        checks["small_grids"] = small_lane()
        progress("checks done")
        if not args.no_modes:
            modes = {}
            for m in ("fp64", "mixed", "bfp", "fp32"):
                if m == args.precision:
                    continue
                ss, el, kt_m, res = lane(probs, m, args.iterations, 1, 1)
                for s in ss:
                    s.close()
                cu, ru = kernel_us(kt_m)
                entry = {"updates_per_s": sum(p.updates for p in probs) / el, "k_col_us": round(cu, 3),
                         "k_row_us": round(ru, 3),
                         "roofline_frac": round(BYTES_PER_UPDATE[m] * npix / ((cu + ru) * 1e-6) / 1e9 / HBM_PEAK_GBPS,
                                                5)}
                if m != "fp64":
                    for i, c in enumerate("xy"):
                        entry[f"E{c}"] = accuracy(pol[c], res[i], refs[c][0], ref_rta[c])
                modes[m] = entry
                progress(f"precision mode {m} done")
    del pol, probs
    other = configs(args.c3_nz, args.c4_count, args.c4_precision, args.configs) if args.configs else None
    serving = served(args.serve_requests, args.precision) if args.serve_requests > 0 else None
    if serving is not None:
        progress("serving done")
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.cpu_reps)
    if cpu is not None:
        progress("CPU baseline done")
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
        "dtype": ARITHMETIC[args.precision],
        "data": "synthetic",
        "config": {
            "workload": f"tilted-interface 3D vector BWPM {args.n}x{args.n}x{args.nz} slabs, "
                        f"{args.iterations} iterations, absorbing zone {args.absorber:g} wavelengths, "
                        f"batch of {args.batch} solves in one lane",
            "n": args.n, "nz": args.nz, "iterations": args.iterations, "batch": args.batch,
            "absorber_wavelengths": args.absorber, "precision": args.precision,
            "streaming_planes": PLANES[args.precision], "arithmetic": ARITHMETIC[args.precision],
            "updates_per_solve": updates // max(args.steps, 1) // args.batch,
        },
        "roofline": roofline,
        "accuracy": headline_acc,
        "checks": checks,
        "precision_modes": modes,
        "configs": other,
        "serving": serving,
        "cpu_baseline": cpu,
        "fft_plans": fft_plans_chosen(),
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
