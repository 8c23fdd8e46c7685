"""Serving loop: kick+drift step requests on resident boxes, continuous batching.

A request asks for ``steps`` fixed-size time steps (kick then drift) of one
box.  The server admits requests in arrival order while the boxes' HBM
footprint fits the budget and a slot is free; an admitted box stays resident
(particles and tree in HBM) until its last step.  Work of different requests
overlaps: while the GPU kicks the boxes that are ready, the host threads build
the trees of the boxes that have just drifted or just arrived, and every box
that is ready when the GPU frees up joins the next batched kick.  A box's
results do not depend on which other boxes share its launches (each box's
kernels run the same operations in the same order), which
``tests/test_serve_gpu.py`` checks bit for bit.

The reference has no serving layer: its driver steps one box
(``kick_step_hierarchical``, src/driver.cpp:382-593: tree_create, kick, drift
per step).  A step here is that sequence with one rung (the time step is
fixed: eta is set so that no particle leaves rung 0).

Executors: ``GpuExecutor`` runs on the device through the C-ABI;
``FakeExecutor`` (tests) simulates the device on the CPU so that admission,
ordering, cancellation and budget rejection are tested without a GPU.
"""
import concurrent.futures as cf
import dataclasses
import threading
import time
from typing import List, Optional

import numpy as np

# HBM bytes per particle of a resident box: x,y,z (12), velocities (12), rung
# (1), saved forces (16), P2P scratch (32), plus tree nodes (64 + 160 + 208 B)
# for ~2 N / 73 nodes and the walk arenas (~1.3 kB per node).
BYTES_PER_PARTICLE = 12 + 12 + 1 + 16 + 32
BYTES_PER_NODE = 64 + 160 + 208 + 1300


def footprint(n, bucket=128):
    nodes = max(1, 2 * n // max(16, bucket // 2))
    return BYTES_PER_PARTICLE * n + BYTES_PER_NODE * nodes


@dataclasses.dataclass
class Request:
    rid: int
    n: int
    steps: int
    seed: int
    clustered: float
    arrival: float  # seconds after the start of the run
    # filled in by the server
    admitted_at: float = -1.0
    finished_at: float = -1.0
    step_latencies: list = dataclasses.field(default_factory=list)
    cancelled: bool = False
    rejected: bool = False
    result: Optional[dict] = None
    # the request's particles (x, y, z, vel), as a client sends them; made on
    # first use when absent (make_particles)
    particles: Optional[tuple] = None


def make_particles(req):
    """The synthetic box of a request: NFW halos plus a uniform background
    (clustered fraction req.clustered), Gaussian velocities, seeded."""
    from . import workloads
    if req.particles is None:
        n = req.n
        x, y, z = workloads.nfw_box(n, req.seed, clustered_fraction=req.clustered, nhalos=max(8, n // 4096))
        vel = workloads.gaussian_velocities(n, req.seed, sigma=1e-3)
        req.particles = (x, y, z, vel)
    return req.particles


def poisson_requests(count, rate, seed=0, nmin=16 ** 3, nmax=192 ** 3, max_steps=10):
    """Requests with exponential inter-arrival times (mean 1/rate), N
    log-uniform in [nmin, nmax], 1..max_steps steps, clustered fraction
    U[0, 0.8]."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), count)).astype(np.int64)
    steps = rng.integers(1, max_steps + 1, count)
    fr = rng.uniform(0.0, 0.8, count)
    return [Request(i, int(ns[i]), int(steps[i]), 5000 + i, float(fr[i]), float(t[i])) for i in range(count)]


class GpuExecutor:
    """Resident boxes on the device (cosmictiger_amd.Box / Batch)."""

    def __init__(self, ct, theta=0.55, bucket=128, dt=1e-3, graph=False):
        self.ct = ct
        self.theta = theta
        self.bucket = bucket
        self.dt = dt
        self.graph = graph
        self.batch = None  # one batch context for every step: its arenas persist
        # seconds spent per operation (summed over the host threads that ran them)
        self.timing = {"create_s": 0.0, "rebuild_s": 0.0, "step_s": 0.0, "steps": 0}
        self._tlock = threading.Lock()

    def _add(self, key, dt):
        with self._tlock:
            self.timing[key] += dt

    def create(self, req):
        x, y, z, vel = make_particles(req)
        t0 = time.perf_counter()
        ct = self.ct
        n = req.n
        box = ct.Box(x, y, z, vel, np.zeros(n, np.int8))
        h = ct.softening(n)
        box.tree_build(self.theta, self.bucket, h)
        self._add("create_s", time.perf_counter() - t0)
        return {"box": box, "h": h, "n": n, "step": 0}

    def rebuild(self, st):
        t0 = time.perf_counter()
        st["box"].tree_build(self.theta, self.bucket, st["h"])
        self._add("rebuild_s", time.perf_counter() - t0)

    def params(self, st):
        p = self.ct.default_kick_params(self.theta, st["h"], 1.0 / st["n"])
        p.t0 = self.dt
        p.first_call = st["step"] == 0
        p.eta = 1e30  # one rung: dt = min(eta sqrt(a h / |g|), t0) = t0 for every particle
        return p

    def step(self, states):
        """One kick + drift of every box in states, together."""
        t0 = time.perf_counter()
        ct = self.ct
        boxes = [s["box"] for s in states]
        if self.batch is None:
            self.batch = ct.Batch(boxes)
        else:
            self.batch.reset(boxes)
        self.batch.step([self.params(s) for s in states], 0, [1.0] * len(states), [self.dt] * len(states), graph=self.graph)
        for s in states:
            s["step"] += 1
        self._add("step_s", time.perf_counter() - t0)
        self._add("steps", 1)

    def step_alone(self, st):
        """The same step with the box's own launches (one box per launch)."""
        t0 = time.perf_counter()
        st["box"].kick(self.params(st))
        st["box"].drift(0, 1.0, self.dt)
        st["step"] += 1
        self._add("step_s", time.perf_counter() - t0)
        self._add("steps", 1)

    def read(self, st):
        return st["box"].read()

    def release(self, st):
        st["box"].close()

    def close(self):
        if self.batch is not None:
            self.batch.close()
            self.batch = None


class FakeExecutor:
    """CPU stand-in for the device: records the calls; a step of a box with n
    particles sleeps n * ns_per_particle seconds.  Used by the CPU tests of the
    scheduler."""

    def __init__(self, ns_per_particle=0.0, fail_on=None):
        self.ns = ns_per_particle
        self.fail_on = fail_on or set()
        self.log = []
        self.lock = threading.Lock()

    def create(self, req):
        if req.rid in self.fail_on:
            raise RuntimeError("create failed")
        with self.lock:
            self.log.append(("create", req.rid))
        return {"rid": req.rid, "n": req.n, "step": 0, "state": np.uint64(req.seed)}

    def rebuild(self, st):
        with self.lock:
            self.log.append(("rebuild", st["rid"]))

    def step(self, states):
        with self.lock:
            self.log.append(("step", tuple(s["rid"] for s in states)))
        time.sleep(sum(s["n"] for s in states) * self.ns * 1e-9)
        for s in states:
            s["state"] = np.uint64((int(s["state"]) * 6364136223846793005 + 1442695040888963407) % (1 << 64))
            s["step"] += 1

    def step_alone(self, st):
        self.step([st])

    def read(self, st):
        return {"state": int(st["state"]), "steps": st["step"]}

    def release(self, st):
        with self.lock:
            self.log.append(("release", st["rid"]))


class Server:
    """Continuous-batching scheduler over an executor.

    hbm_budget: bytes of resident boxes admitted at once (footprint());
    max_slots: boxes resident at once; host_threads: concurrent tree builds.
    A request whose own footprint exceeds the budget is rejected."""

    def __init__(self, executor, hbm_budget, max_slots=64, host_threads=8, batched=True):
        self.ex = executor
        self.budget = hbm_budget
        self.max_slots = max_slots
        self.pool = cf.ThreadPoolExecutor(max_workers=host_threads)
        self.batched = batched
        self._cancel = set()
        self.max_used = 0  # largest admitted footprint seen in the last run
        self.max_batch = 0  # most boxes stepped together in the last run

    def cancel(self, rid):
        """Cancel a request: it takes no further step and its box is freed."""
        self._cancel.add(rid)

    def run(self, requests: List[Request], clock=time.perf_counter):
        reqs = sorted(requests, key=lambda r: r.arrival)
        t0 = clock()
        now = lambda: clock() - t0
        waiting = list(reqs)
        pending = []
        active = {}  # rid -> dict(req, state, future, ready_at)
        used = 0
        done = []
        while waiting or pending or active:
            t = now()
            while waiting and waiting[0].arrival <= t:
                r = waiting.pop(0)
                if footprint(r.n) > self.budget:
                    r.rejected = True
                    done.append(r)
                else:
                    pending.append(r)
            # admission in arrival order while memory and slots allow
            while pending and len(active) < self.max_slots and used + footprint(pending[0].n) <= self.budget:
                r = pending.pop(0)
                r.admitted_at = now()
                used += footprint(r.n)
                self.max_used = max(self.max_used, used)
                active[r.rid] = {"req": r, "state": None, "future": self.pool.submit(self.ex.create, r), "ready_at": r.arrival}
            # cancellations
            for rid in list(active):
                if rid in self._cancel:
                    a = active[rid]
                    st = a["future"].result() if a["future"] is not None else a["state"]
                    if st is not None:
                        self.ex.release(st)
                    a["req"].cancelled = True
                    a["req"].finished_at = now()
                    used -= footprint(a["req"].n)
                    done.append(a["req"])
                    del active[rid]
            ready = []
            for rid, a in active.items():
                f = a["future"]
                if f is not None and f.done():
                    res = f.result()
                    if a["state"] is None:
                        a["state"] = res
                    a["future"] = None
                if a["future"] is None:
                    ready.append(a)
            if not ready:
                futs = [a["future"] for a in active.values() if a["future"] is not None]
                if futs:
                    cf.wait(futs, timeout=0.005, return_when=cf.FIRST_COMPLETED)
                elif waiting and not pending:
                    time.sleep(max(0.0, min(0.005, waiting[0].arrival - now())))
                else:
                    time.sleep(0.0005)
                continue
            self.max_batch = max(self.max_batch, len(ready))
            if self.batched:
                self.ex.step([a["state"] for a in ready])
            else:
                # one box per launch, in arrival order of the ready boxes
                for a in sorted(ready, key=lambda q: q["req"].arrival):
                    self.ex.step_alone(a["state"])
            t_end = now()
            for a in ready:
                r = a["req"]
                r.step_latencies.append(t_end - a["ready_at"])
                if len(r.step_latencies) >= r.steps:
                    r.result = self.ex.read(a["state"])
                    self.ex.release(a["state"])
                    r.finished_at = t_end
                    used -= footprint(r.n)
                    done.append(r)
                    del active[r.rid]
                else:
                    a["ready_at"] = t_end
                    a["future"] = self.pool.submit(self.ex.rebuild, a["state"])
        wall = now()
        return summarize(done, wall)

    def close(self):
        self.pool.shutdown(wait=True)


def summarize(done, wall):
    served = [r for r in done if not r.rejected and not r.cancelled]
    lat = [x for r in served for x in r.step_latencies]
    psteps = sum(r.n * len(r.step_latencies) for r in served)
    return {"requests": len(done), "served": len(served), "rejected": sum(r.rejected for r in done),
            "cancelled": sum(r.cancelled for r in done), "steps": len(lat), "particle_steps": psteps, "wall_s": wall,
            "throughput": psteps / wall if wall > 0 else 0.0,
            "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)) if lat else None,
                           "p99": 1e3 * float(np.percentile(lat, 99)) if lat else None}}
