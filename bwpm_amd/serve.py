"""Serving: many BWPM solves of uneven sizes on one MI355X with continuous batching.

`Server` wraps the library's request queue (`_core.GpuQueue`). Requests are admitted in arrival
order while their device footprint fits the HBM budget. Admitted solves of one (precision, grid)
step together in a continuous lane, one column and one row launch per slab step for the whole lane.
A request that arrives while others are on the GPU joins a free slot between chunks of slab steps,
so its work overlaps with theirs. Every result is bit-identical to the same solve run alone.
Latency is measured from submit to result on the host; the queue reports p50/p99 and throughput in
grid-point·slab updates/s.

`served_request(seed)` draws one request of the served mix, and `run_stream` submits requests at
Poisson arrival times and collects the results. The mix has 2D solves on 512 … 4096 points and 3D
vector solves on 128² … 512² with 200 … 1024 slabs. Geometries are planar, tilted, a silver-like
layer and δn volumes; their label volumes are made inside the library from a `Geometry` and
their δn on the device, so a request crosses the C ABI as a few hundred bytes plus its field.
"""

from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field
from typing import Any, Callable, Iterator

import numpy as np

from bwpm_amd import _core, workloads
from bwpm_amd.specs import Geometry, Layer

SILVER = complex(0.13, 4.0)


class Server:
    """A request queue on one device: `submit(problem)` → ticket, `result(ticket)` → dict."""

    def __init__(self, device: int = 0, hbm_budget: float | None = None, slots: int = 16, chunk: int = 64,
                 precision: str = "fp64") -> None:
        self.precision = precision
        self._q = _core.GpuQueue(device, hbm_budget, slots, chunk)

    def submit(self, p: workloads.Problem, precision: str | None = None) -> int:
        return self._q.submit(p.ny, p.nx, p.dy, p.dx, p.dz, p.k0, p.indices, p.labels, p.index_before, p.mask,
                              p.field, p.iterations, fluctuation=p.fluctuation,
                              precision=precision or self.precision)

    def state(self, ticket: int) -> str:
        return self._q.state(ticket)

    def wait(self, ticket: int, timeout: float | None = None) -> str:
        return self._q.wait(ticket, timeout)

    def result(self, ticket: int, timeout: float | None = None) -> dict[str, Any]:
        """Waits for the ticket and returns its result (RuntimeError/ValueError if it failed)."""
        self._q.wait(ticket, timeout)
        return self._q.result(ticket)

    def stats(self) -> dict[str, float]:
        return self._q.stats()

    def close(self) -> None:
        self._q.close()

    def __enter__(self) -> "Server":
        return self

    def __exit__(self, *exc: Any) -> None:
        self.close()


# ---- the served mix --------------------------------------------------------------------------

SIZES_2D = (512, 576, 768, 1024, 1152, 1536, 2048, 3072, 4096)  # 2^a·3^b
SIZES_3D = (128, 192, 256, 384, 512)


def planar_3d(n: int, nz: int, iterations: int, wavelength_um: float, n1: complex, n2: complex,
              layer: tuple[int, int] | None = None, polarisation: str = "x", geometry: bool = True) -> workloads.Problem:
    """Planar interface n1 → n2 at mid-depth, or a layer of n2 in n1 in slabs layer[0] … layer[1] − 1.
    With `geometry` the labels are made inside the library (no (nz, n, n) array on the host)."""
    p = workloads.tilted_3d(n=n, nz=nz, iterations=iterations, wavelength_um=wavelength_um, tilt_deg=0.0,
                            n1=n1, n2=n2, polarisation=polarisation, geometry=geometry)
    if layer is not None:
        if geometry:
            p.labels = Geometry(nz, (Layer(1, layer[0] * p.dz, layer[1] * p.dz),))
        else:
            lab = np.zeros((nz, n, n), dtype=np.int32)
            lab[layer[0]:layer[1]] = 1
            p.labels = lab
    return p


def served_request(seed: int, max_n: int = 512, max_nz: int = 1024, born_n: int | None = None) -> workloads.Problem:
    """One request of the served mix, drawn from `seed` (reproducible). δn requests have the sizes of
    the other 3D requests (n² up to max_n², 200 … max_nz slabs, 1 … 3 iterations); `born_n` fixes
    them to born_n² × 50 … 128 slabs × 1 iteration instead (small test mixes)."""
    rng = np.random.default_rng(seed)
    kind = rng.choice(["2d", "planar", "tilted", "silver", "born"], p=[0.2, 0.25, 0.3, 0.15, 0.1])
    its = int(rng.integers(1, 4))
    if kind == "2d":
        ny = int(rng.choice(SIZES_2D))
        nz = int(rng.integers(200, 1025))
        p = workloads.plain_2d(ny=ny, nz=nz, iterations=its)
        p.name = f"served_2d_{ny}x{nz}"
        return p
    sizes = [s for s in SIZES_3D if s <= max_n]
    n = int(rng.choice(sizes))
    nz = int(rng.integers(200, max_nz + 1))
    lam = float(rng.uniform(0.6, 1.6))
    pol = "x" if rng.random() < 0.5 else "y"
    if kind == "planar":
        p = planar_3d(n, nz, its, lam, 1.0, float(rng.uniform(1.3, 1.7)), polarisation=pol)
    elif kind == "tilted":
        p = workloads.tilted_3d(n=n, nz=nz, iterations=its, wavelength_um=lam, tilt_deg=float(rng.uniform(2, 15)),
                                n2=float(rng.uniform(1.3, 1.7)), polarisation=pol, geometry=True)
    elif kind == "silver":
        z0 = int(rng.integers(nz // 4, nz // 2))
        p = planar_3d(n, nz, its, lam, 1.0, SILVER, layer=(z0, z0 + int(rng.integers(2, 8))), polarisation=pol)
    elif born_n is None:
        p = workloads.born_3d_synth(n=n, nz=nz, iterations=its, sigma=float(rng.uniform(0.005, 0.02)),
                                    seed=int(rng.integers(1 << 30)), wavelength_um=lam, step_um=lam / 6,
                                    dz_um=lam / 4)
    else:
        nzb = int(rng.integers(50, 129))
        p = workloads.born_3d_synth(n=born_n, nz=nzb, iterations=1, sigma=float(rng.uniform(0.005, 0.02)),
                                    seed=int(rng.integers(1 << 30)))
    p.name = f"served_{kind}_{p.ny}x{p.nx}x{p.nz}"
    return p


@dataclass
class StreamReport:
    """Outcome of a served stream: the queue's statistics (latency submit → result on the host,
    p50/p99, throughput) and the failed requests."""

    submitted: int
    offered_rate_hz: float
    wall_s: float
    stats: dict[str, float]
    failures: list[str] = field(default_factory=list)
    results: dict[int, dict[str, Any]] = field(default_factory=dict)


def run_stream(server: Server, requests: Iterator[workloads.Problem] | list[workloads.Problem], rate_hz: float,
               seed: int = 0, keep_results: bool = False,
               on_result: Callable[[int, workloads.Problem, dict[str, Any]], None] | None = None) -> StreamReport:
    """Submits `requests` at Poisson arrival times of mean rate `rate_hz` (seeded) while a collector
    thread takes results as they finish; returns when every request has finished."""
    rng = np.random.default_rng(seed)
    tickets: list[tuple[int, workloads.Problem]] = []
    lock = threading.Lock()
    done = threading.Event()
    report = StreamReport(0, rate_hz, 0.0, {})

    def collect() -> None:
        i = 0
        while True:
            with lock:
                item = tickets[i] if i < len(tickets) else None
            if item is None:
                if done.is_set():
                    with lock:
                        if i >= len(tickets):
                            return
                    continue
                time.sleep(0.001)
                continue
            t, prob = item
            try:
                res = server.result(t)
                if on_result:
                    on_result(t, prob, res)
                if keep_results:
                    report.results[t] = res
            except (RuntimeError, ValueError) as err:
                report.failures.append(f"{prob.name}: {err}")
            i += 1

    collector = threading.Thread(target=collect, daemon=True)
    t0 = time.perf_counter()
    collector.start()
    next_at = t0
    for prob in requests:
        now = time.perf_counter()
        if next_at > now:
            time.sleep(next_at - now)
        t = server.submit(prob)
        with lock:
            tickets.append((t, prob))
        report.submitted += 1
        next_at += float(rng.exponential(1.0 / rate_hz))
    done.set()
    collector.join()
    report.wall_s = time.perf_counter() - t0
    report.stats = server.stats()
    return report


__all__ = ["SILVER", "Server", "StreamReport", "planar_3d", "run_stream", "served_request"]
