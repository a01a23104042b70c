"""Ensemble serving loop over the device pools (gdf_pool): a stream of independent simulation
requests, each admitted when its class's batch has room for its particles (the HBM footprint
limit), advanced with the other resident requests of its class one batched step at a time, and
retired after its steps.  One pool per configuration class (the periodic hydro box, the
self-gravitating cold sphere and the Plummer N-body cluster); each pool is driven by its own
thread, so the classes' steps overlap on the device (every pool has its own HIP stream and the
C ABI releases the GIL).

A request's per-step latency is the wall time of the batch step that advanced it; queueing
before admission is reported separately.
"""
from __future__ import annotations

import dataclasses
import threading
import time

import numpy as np

from . import ic as ICS
from .gpu import GpuPool, GpuSim

# device bytes per resident particle, rounded up from the pool context's allocations per
# particle of capacity: 23 persistent FP64 fields and their permutation targets (368 B), 25
# derived fields (200 B), integer state and ids (64 B), packed records and export scratch
# (96 B), tree arrays over 2N nodes (~400 B), and the two candidate sets at their initial 16
# batches of 64 per walk group (~512 B): ~1.65 KB; the admission budget divided by this gives
# a pool's particle limit
BYTES_PER_PARTICLE = 2048

CLASSES = ("box", "sphere", "plummer")


def make_request_ic(kind: str, n: int, seed: int) -> dict:
    """Initial conditions of one request: the three generators of the workload, n particles
    (the lattice box takes the nearest cube)."""
    if kind == "box":
        side = max(8, int(round(n ** (1.0 / 3.0))))
        return ICS.jittered_lattice_box(n_side=side, seed=seed)
    if kind == "sphere":
        return ICS.cold_sphere(N=n, seed=seed)
    if kind == "plummer":
        return ICS.plummer_sphere(N=n, seed=seed)
    raise ValueError(kind)


def class_params(kind: str) -> dict:
    """The configuration class of a generator (its parameters do not depend on n or the seed)."""
    return make_request_ic(kind, 512, 0)["params"]


@dataclasses.dataclass
class Request:
    rid: int
    kind: str
    n: int
    seed: int
    arrival: float
    nsteps: int
    ic: dict = None
    pool_id: int = -1
    admitted: float = -1.0
    finished: float = -1.0
    step_latency: list = dataclasses.field(default_factory=list)


def request_stream(nreq=256, rate=20.0, nmin=4096, nmax=262144, nsteps=20, seed=5):
    """Poisson arrivals (mean `rate` per second), sizes log-uniform in [nmin, nmax], the three
    generators in rotation, per-request seeds."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, nreq))
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), nreq)).astype(int)
    return [Request(k, CLASSES[k % 3], int(sizes[k]), int(seed * 100003 + k), float(t[k]), nsteps)
            for k in range(nreq)]


class Server:
    """Drives one pool per class over a request stream in real time."""

    def __init__(self, budget_bytes=96 << 30, precision="mixed"):
        self.precision = precision
        self.max_particles = int(budget_bytes // BYTES_PER_PARTICLE // len(CLASSES))
        self.pools = {k: GpuPool(class_params(k), precision, self.max_particles) for k in CLASSES}
        self.lock = threading.Lock()
        self.psteps = {k: 0 for k in CLASSES}
        self.error = None

    def close(self):
        for p in self.pools.values():
            p.close()

    def _run_class(self, kind, reqs, t0, keep_results):
        pool = self.pools[kind]
        queue = sorted(reqs, key=lambda q: q.arrival)
        waiting, resident = [], {}
        try:
            while queue or waiting or resident:
                now = time.perf_counter() - t0
                while queue and queue[0].arrival <= now:
                    waiting.append(queue.pop(0))
                # admission in arrival order while the batch has room
                while waiting:
                    q = waiting[0]
                    g = q.ic if q.ic is not None else make_request_ic(q.kind, q.n, q.seed)
                    pid = pool.admit(g["r"], g["v"], g["m"], g["u"], q.nsteps)
                    if pid is None:
                        break
                    q.pool_id = pid
                    q.admitted = time.perf_counter() - t0
                    q.n = g["r"].shape[0]
                    resident[pid] = q
                    waiting.pop(0)
                if not resident:
                    if queue:
                        time.sleep(max(0.0, min(queue[0].arrival - now, 0.05)))
                    continue
                ts = time.perf_counter()
                _, ps = pool.step()
                dt = time.perf_counter() - ts
                self.psteps[kind] += ps
                done = []
                for pid, q in resident.items():
                    q.step_latency.append(dt)
                    if len(q.step_latency) >= q.nsteps:
                        q.finished = time.perf_counter() - t0
                        if keep_results:
                            q.ic = {f: pool.get(pid, f) for f in ("r", "v", "a", "h", "u")}
                        pool.release(pid)
                        done.append(pid)
                for pid in done:
                    del resident[pid]
        except Exception as e:          # surfaced by run()
            self.error = e

    def run(self, reqs, keep_results=False):
        """Serve the requests in real time from now; returns the summary."""
        by = {k: [q for q in reqs if q.kind == k] for k in CLASSES}
        t0 = time.perf_counter()
        th = [threading.Thread(target=self._run_class, args=(k, by[k], t0, keep_results))
              for k in CLASSES if by[k]]
        for t in th:
            t.start()
        for t in th:
            t.join()
        wall = time.perf_counter() - t0
        if self.error is not None:
            raise self.error
        return summarize(reqs, wall, sum(self.psteps.values()))


def summarize(reqs, wall, psteps):
    lat = np.concatenate([np.asarray(q.step_latency) for q in reqs if q.step_latency])
    wait = np.asarray([q.admitted - q.arrival for q in reqs if q.admitted >= 0])
    return {"requests": len(reqs), "particle_steps": int(psteps), "wall_s": wall,
            "value": psteps / wall,
            "step_latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)),
                                "p99": 1e3 * float(np.percentile(lat, 99)),
                                "max": 1e3 * float(lat.max())},
            "admission_wait_ms": {"p50": 1e3 * float(np.percentile(wait, 50)),
                                  "p99": 1e3 * float(np.percentile(wait, 99))}}


def run_sequential(reqs, precision="mixed"):
    """The same requests one after another, each alone on the device (one reused context per
    class): end-to-end time per request (set-up and its steps) and per-step latencies."""
    ctx = {}
    psteps, busy, steps_only, lat = 0, 0.0, 0.0, []
    for q in reqs:
        g = make_request_ic(q.kind, q.n, q.seed)
        if q.kind not in ctx:
            ctx[q.kind] = GpuSim(g["params"], precision=precision)
        s = ctx[q.kind]
        t0 = time.perf_counter()
        s.set_particles(g["r"], g["v"], g["m"], g["u"])
        s.setup()
        for _ in range(q.nsteps):
            ts = time.perf_counter()
            s.step(1)
            d = time.perf_counter() - ts
            steps_only += d
            lat.append(d)
        busy += time.perf_counter() - t0
        psteps += g["r"].shape[0] * q.nsteps
    for s in ctx.values():
        s.close()
    lat = np.asarray(lat)
    return {"requests": len(reqs), "particle_steps": int(psteps), "wall_s": busy,
            "value": psteps / busy, "steps_only_value": psteps / steps_only,
            "step_latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)),
                                "p99": 1e3 * float(np.percentile(lat, 99))}}
