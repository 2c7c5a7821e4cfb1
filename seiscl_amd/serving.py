"""Serving: continuous batching of modeling and gradient requests on one GPU.

A request is one modeling job (model, shot, constants) or one gradient job
(the same plus observed seismograms).  Requests arrive over time.  Modeling
jobs that the batch kernels cover (2D, FP32, L=0, no free surface) are
admitted into a batch (seis_batch) while slots and the HBM budget allow;
every admitted job advances together -- one launch per kernel class and time
step for the whole batch -- and returns its seismograms as soon as it has run
its last step, freeing its slot for the next arrival.  Gradient jobs, and
modeling jobs the batch does not cover, run one after another on a second
context of the same GPU (its own streams), concurrently with the batch.  The
reference has no counterpart: it runs one shot after another
(src/time_stepping.c:747), which is the plain path timed here for comparison
(one model, one launch sequence per shot, first come first served).

Latency is arrival -> result on the host; throughput is shots per hour over
the run.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from .engine import Batch, Context, Model


@dataclass
class Request:
    rid: int
    csts: dict
    params: dict
    shot: tuple
    arrival: float = 0.0          # seconds after the start of the run
    start: float = -1.0           # admitted
    finish: float = -1.0          # result on the host
    out: list = field(default_factory=list)
    din: list | None = None       # observed seismograms: a gradient request
    grad: dict = field(default_factory=dict)   # its gradient (Model.grad_read)
    misfit: tuple = ()            # its (rms, rmsnorm)

    @property
    def kind(self) -> str:
        return "forward" if self.din is None else "gradient"

    def cell_updates(self) -> int:
        return int(np.prod(self.csts["N"])) * int(self.csts["NT"])


def step_bytes(csts: dict) -> int:
    """Bytes one time step of a request touches (wavefields, coefficients,
    CPML memory): its share of the last-level cache while it is active."""
    NZ, NX = (int(v) for v in csts["N"])
    H = int(csts["FDORDER"]) // 2
    cells = (NZ + 2 * H + 64) * (NX + 2 * H)
    nf = 3 if csts["ND"] == 22 else 5
    npar = 3 if csts["ND"] == 22 else 5
    nab = int(csts.get("nab", 16))
    psi = 8 * 2 * nab * (NZ + NX)
    return 4 * (cells * (nf + npar) + psi)


def estimate_bytes(csts: dict) -> int:
    """HBM a request needs (wavefields, coefficients, CPML memory, shot
    buffers), from the sizes alone (the model's own footprint is known only
    once it exists)."""
    NX = int(csts["N"][1])
    return step_bytes(csts) + 8 * int(csts["NT"]) * (NX // 2 + 8)


def poisson_arrivals(n: int, rate: float, seed: int = 40) -> np.ndarray:
    """Arrival times of a Poisson process of the given rate (per second)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n)) - 0.0


def _wait_until(t0: float, t: float):
    dt = t - (time.perf_counter() - t0)
    if dt > 0:
        time.sleep(dt)


def batchable(csts: dict) -> bool:
    """Whether the batch kernels cover a modeling request (seis_batch_add)."""
    return (int(csts["ND"]) in (2, 22) and int(csts.get("L", 0)) == 0
            and not int(csts.get("freesurf", 0)) and int(csts.get("FP16", 0)) == 0)


def run_alone(ctx: Context, r: Request):
    """One request on its own model: seismograms, and for a gradient request
    its gradient and misfit (src/time_stepping.c:747-866 for one shot)."""
    m = Model(ctx, r.params, **r.csts)
    try:
        if r.din is None:
            r.out = m.forward(*r.shot)
        else:
            m.grad_zero()
            r.out, r.misfit = m.gradient(*r.shot, din=r.din)
            r.grad = m.grad_read()
    finally:
        m.close()


def serve_plain(ctx: Context, requests):
    """One job at a time in arrival order (the reference's execution model)."""
    t0 = time.perf_counter()
    for r in sorted(requests, key=lambda q: q.arrival):
        _wait_until(t0, r.arrival)
        r.start = time.perf_counter() - t0
        run_alone(ctx, r)
        r.finish = time.perf_counter() - t0
    return requests


def serve_batched(ctx: Context, requests, max_slots: int = 48, hbm_budget: float = 64e9,
                  chunk: int = 64, prefetch: int = 4, cache_budget: float = 192e6,
                  min_slots: int = 8, setup_threads: int = 4):
    """Continuous batching.  Set-up threads (setup_threads of them) build
    each admitted request's model (host transforms, uploads on the context's
    transfer stream) while the GPU steps the batch; the main loop moves ready
    models into free slots, queues the next steps, and only then copies back
    the jobs that finished in the previous steps, so the GPU always has work
    queued.  Gradient requests and modeling requests outside the batch
    kernels go, in arrival order, to a worker thread that runs them one at a
    time on a second context (run_alone), beside the batch.

    Admission to set-up is bounded by HBM (hbm_budget); activation in the
    batch by the bytes a step touches: ready models join while the active
    ones' step_bytes stay within cache_budget (three quarters of MI355X's
    256 MB last-level cache, so that a batch step runs from that cache
    rather than from HBM), but never fewer than min_slots are active while
    models wait, nor more than max_slots."""
    import queue
    import threading

    batch = Batch(ctx, max_slots)
    pending = deque(sorted(requests, key=lambda q: q.arrival))
    todo, ready = queue.Queue(), queue.Queue()
    errors = []
    t0 = time.perf_counter()
    # the worker of the requests that run alone, on a context of its own
    alone_q, alone_done = queue.Queue(), queue.Queue()
    actx = Context(ctx.device) if any(
        r.din is not None or not batchable(r.csts) for r in pending) else None

    def alone():
        while True:
            r = alone_q.get()
            if r is None:
                return
            try:
                run_alone(actx, r)
            except Exception as e:          # reported by the main loop
                errors.append(e)
            r.finish = time.perf_counter() - t0
            alone_done.put(r)

    def setup():
        while True:
            r = todo.get()
            if r is None:
                return
            try:
                m = Model(ctx, r.params, **r.csts)
                m.bind_shot(*r.shot, init=False)
                ready.put((r, m))
            except Exception as e:          # reported by the main loop
                errors.append(e)
                ready.put((r, None))

    workers = [threading.Thread(target=setup, daemon=True) for _ in range(max(1, setup_threads))]
    if actx is not None:
        workers.append(threading.Thread(target=alone, daemon=True))
    for w in workers:
        w.start()
    active = {}                 # slot -> (request, model, bytes)
    done_prev = []              # slots that ran their last step in the previous round
    inflight = 0                # requests handed to the set-up thread, not yet active
    solo = 0                    # requests handed to the alone worker, not yet done
    used = 0.0                  # HBM of admitted requests
    hot = 0.0                   # step bytes of the active requests
    held = None                 # a ready model waiting for cache room
    try:
        while pending or inflight or active or solo:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now and (
                    pending[0].din is not None or not batchable(pending[0].csts)):
                r = pending.popleft()
                r.start = now
                solo += 1
                alone_q.put(r)
            while not alone_done.empty():
                alone_done.get()
                solo -= 1
            if errors:
                raise errors[0]
            while (pending and pending[0].arrival <= now
                   and pending[0].din is None and batchable(pending[0].csts)
                   and len(active) + inflight < max_slots + prefetch):
                need = estimate_bytes(pending[0].csts)
                if (active or inflight) and used + need > hbm_budget:
                    break
                r = pending.popleft()
                r.start = now
                used += need
                inflight += 1
                todo.put(r)
            while len(active) < max_slots:
                if held is None:
                    try:
                        held = ready.get_nowait()
                    except queue.Empty:
                        break
                r, m = held
                if m is None:
                    raise errors[0]
                if len(active) >= max(1, min_slots) and hot + step_bytes(r.csts) > cache_budget:
                    break
                held = None
                inflight -= 1
                hot += step_bytes(r.csts)
                active[batch.add(m)] = (r, m, estimate_bytes(r.csts))
            nstep, min_left = batch.status()
            if nstep:
                batch.step(max(1, min(min_left, chunk)))
            # copy back what finished before the steps just queued
            for slot in done_prev:
                r, m, need = active.pop(slot)
                r.out = m.read_shot()
                r.finish = time.perf_counter() - t0
                batch.remove(slot)
                m.close()
                used -= need
                hot -= step_bytes(r.csts)
            done_prev = [s for s in active if batch.left(s) == 0]
            if not nstep and not done_prev:
                # nothing in the batch: wait for a model, the next arrival or
                # the alone worker
                if held is None and inflight:
                    held = ready.get()
                elif solo:
                    wait = None
                    if pending:
                        wait = max(0.0, pending[0].arrival - (time.perf_counter() - t0))
                    try:
                        alone_done.get(timeout=wait)
                        solo -= 1
                    except queue.Empty:
                        pass
                elif pending:
                    _wait_until(t0, pending[0].arrival)
    finally:
        for _ in range(max(1, setup_threads)):
            todo.put(None)
        alone_q.put(None)
        for w in workers:
            w.join()
        if held is not None and held[1] is not None:
            held[1].close()
        while not ready.empty():
            _, m = ready.get()
            if m is not None:
                m.close()
        for r, m, _ in active.values():
            m.close()
        batch.close()
        if actx is not None:
            actx.close()
    return requests


def summarize(requests) -> dict:
    lat = np.array([r.finish - r.arrival for r in requests])
    span = max(r.finish for r in requests) - min(r.arrival for r in requests)
    cu = sum(r.cell_updates() for r in requests)
    return dict(shots_per_hour=len(requests) / span * 3600.0, span_s=span,
                p50_latency_s=float(np.percentile(lat, 50)),
                p99_latency_s=float(np.percentile(lat, 99)),
                gcell_s=cu / span / 1e9)


def make_requests(workload, arrivals=None, observed=None):
    """Requests from (csts, params, shot) jobs; observed[i] (seismograms, or
    None) makes job i a gradient request."""
    reqs = []
    for i, (csts, params, shot) in enumerate(workload):
        reqs.append(Request(i, csts, params, shot,
                            arrival=float(arrivals[i]) if arrivals is not None else 0.0,
                            din=observed[i] if observed is not None else None))
    return reqs
