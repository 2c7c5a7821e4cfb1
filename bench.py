#!/usr/bin/env python3
"""Benchmark: 3D viscoelastic forward modeling on one MI355X.

Workload: one shot on a 256^3 grid (dh 10 m, dt 0.8 ms, FD order 4, CPML 16,
free surface, L=2 Maxwell bodies with Q=50 at 5-25 Hz, layered model with 3%
smooth perturbations, 12 Hz Ricker explosion, 64 x 64 surface receivers),
FP32.  A "step" is one time step of the finite-difference loop over the
whole grid: sources, velocity update, stress update, free surface and
receiver sampling.  W warm-up steps run untimed, then K steps are timed with
the wavefields resident in HBM; the shot then runs on to its last step
(NT = 1500) untimed, so that its seismograms can be compared.

value = grid cells x K / timed seconds  (Gcell-updates/s).
roofline: the stress update (the dominant kernel), algorithmic bytes per
launch = 196 B per cell (3 velocities read, 6 stresses and 12 memory
variables read-modify-write, 10 coefficient arrays read) / its average
duration from HIP events on the compute stream.
roofline.traffic: HBM bytes per stress-update launch, measured by two
rocprofv3 --pmc child runs of this script (FETCH_SIZE, then WRITE_SIZE),
each scaled by what the counter reads for a 512 MiB device copy in the same
run; they run before this process touches the GPU (--no-traffic skips
them).
cpu_baseline: the C oracle on the host cores, same grid, a bounded number of
steps.
config.elastic: the same shot without attenuation (L=0), the 3D elastic
case of the metric, timed the same way with its own roofline fractions.
config.fp16_storage / config.fp64: the same shot run with FP16 wavefield and
memory-variable storage (FP32 arithmetic) and in FP64; seismogram errors are
relative L2 norms over the whole shot against the FP64 run (skip with
--no-variants).
config.serving: a bounded run of the serving workload (64 seeded 2D
requests): capacity of the plain and batched paths and a Poisson load at
the plain path's capacity, shots/h with p50/p99 latency (--no-serving skips
it; --workload serving runs the full 256 requests).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s & % HBM roofline, 3D elastic; shots/hr + p99 latency batched"
PEAK_HBM_GBS = 8000.0
BYTES_PER_CELL_S = 196       # stress update, 3D visco L=2 FP32
BYTES_PER_CELL_V = 60        # velocity update, 3D FP32
BYTES_PER_CELL_FP16 = 154    # whole step, FP16 fields and memory, FP32 coefficients
BYTES_PER_CELL_S_EL = 80   # stress update, 3D elastic FP32
BYTES_PER_CELL_EL = 140     # whole step, 3D elastic FP32
COPY_BYTES = 512 * 2**20    # calibration copy of the HBM counters


def traffic_probe(args):
    """Child run under rocprofv3 --pmc: a 512 MiB device copy (counter
    calibration), then a few steps of the benchmark shot."""
    import torch
    x = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    del x, y
    log("probe: calibration copy done")
    from seiscl_amd import Context, Model
    from seiscl_amd.workloads import config2
    csts, params, shots = config2(n=args.n, NT=12, elastic=args.traffic_probe == "elastic")
    ctx = Context(0)
    m = Model(ctx, params, **csts)
    m.bind_shot(*shots[0])
    m.run_steps(0, 12)
    ctx.sync()
    m.close()
    ctx.close()
    log("probe: 12 steps done")


def measure_traffic(n, elastic=False, timeout=170):
    """HBM bytes per launch of the stress and velocity updates from two
    rocprofv3 --pmc child runs (FETCH_SIZE, WRITE_SIZE), each counter scaled
    by its reading for the 512 MiB copy of the same run.  Returns
    ({kernel: bytes}, note) or (None, reason)."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None, "rocprofv3 not found"
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="seis_pmc_")
        cmd = [exe, "--pmc", counter, "-f", "csv", "-d", d, "-o", "run", "--", sys.executable,
               os.path.abspath(__file__), "--traffic-probe", "elastic" if elastic else "visco",
               "--n", str(n)]
        log(f"traffic: {counter} pass ({'elastic' if elastic else 'viscoelastic'})")
        try:
            r = subprocess.run(cmd, stdout=sys.stderr, stderr=sys.stderr, timeout=timeout,
                               cwd=ROOT, env=dict(os.environ, TMPDIR="/tmp"))
        except subprocess.TimeoutExpired:
            return None, f"rocprofv3 {counter} pass timed out"
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if r.returncode != 0 and not files:
            return None, f"rocprofv3 {counter} pass failed (rc {r.returncode})"
        per = {}
        for f in files:
            for row in csv.DictReader(open(f)):
                if row.get("Counter_Name") != counter:
                    continue
                per.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
        shutil.rmtree(d, ignore_errors=True)
        copy = [v for k, vals in per.items() if "copyBuffer" in k for v in vals]
        copy = [v for v in copy if v > 0.25 * COPY_BYTES / 1024]
        if not copy:
            return None, f"no calibration copy in the {counter} pass"
        factor = COPY_BYTES / (1024.0 * max(copy))
        for name, tag in (("k_s3d", "update_s"), ("k_v3d", "update_v")):
            vals = [v for k, vs in per.items() if name in k for v in vs]
            if vals:
                res.setdefault(tag, {})[counter] = (1024.0 * sum(vals) / len(vals) * factor,
                                                     factor)
    out = {k: int(v["FETCH_SIZE"][0] + v["WRITE_SIZE"][0]) for k, v in res.items()
           if "FETCH_SIZE" in v and "WRITE_SIZE" in v}
    if "update_s" not in out:
        return None, "stress-update counters missing"
    f_fetch = next(iter(res.values()))["FETCH_SIZE"][1]
    f_write = next(iter(res.values()))["WRITE_SIZE"][1]
    return out, f"FETCH_SIZE x {f_fetch:.3f} + WRITE_SIZE x {f_write:.3f} (512 MiB copy)"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(csts, params, shot, steps, threads):
    import oracle
    sp, s, rp = shot
    c = dict(csts, NT=steps)
    t0 = time.time()
    oracle.forward(params, sp, s[:steps], rp, nthreads=threads, **c)
    wall = time.time() - t0
    loop = oracle.last_loop_seconds()
    n = int(np.prod(csts["N"]))
    return dict(value=n * steps / loop / 1e9, unit="Gcell/s", cores=threads, kind="port",
                sample=(f"{steps} time steps of the same {tuple(int(v) for v in csts['N'])} grid "
                        f"(loop {loop:.1f} s, {wall:.1f} s with set-up), "
                        "C oracle with OpenMP"))


def serving_main(args):
    """Serving workload: seeded 2D acoustic/elastic requests of uneven sizes
    and lengths, Poisson arrivals at 0.5x, 1x and 2x the plain path's
    measured capacity; continuous batching against one shot at a time."""
    from seiscl_amd import Context
    from seiscl_amd.serving import (make_requests, poisson_arrivals, serve_batched,
                                    serve_plain, summarize)
    from seiscl_amd.workloads import config4
    work = config4(args.requests)
    ctx = Context(0)
    # warm-up: one small job of each kind compiles nothing but touches every
    # kernel once
    serve_plain(ctx, make_requests(work[:4]))
    cap_plain = summarize(serve_plain(ctx, make_requests(work)))
    cap_batch = summarize(serve_batched(ctx, make_requests(work), max_slots=args.slots))
    log(f"capacity: plain {cap_plain['shots_per_hour']:.0f} shots/h, "
        f"batched {cap_batch['shots_per_hour']:.0f} shots/h")
    rate0 = cap_plain["shots_per_hour"] / 3600.0
    loads = {}
    for f in (0.5, 1.0, 2.0):
        arr = poisson_arrivals(len(work), f * rate0)
        p = summarize(serve_plain(ctx, make_requests(work, arr)))
        b = summarize(serve_batched(ctx, make_requests(work, arr), max_slots=args.slots))
        loads[f"{f}x"] = {"plain": p, "batched": b}
        log(f"load {f}x: plain {p['shots_per_hour']:.0f}/h p99 {p['p99_latency_s']:.2f} s; "
            f"batched {b['shots_per_hour']:.0f}/h p99 {b['p99_latency_s']:.2f} s")
    mixed = serving_mixed(ctx, work, args.grad_share, args.slots) if args.grad_share > 0 else None
    ctx.close()
    value = cap_batch["shots_per_hour"]
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "shots/h", "n_gpus": 1,
        "steps": len(work), "warmup": 4,
        "ms_per_step": round(cap_batch["span_s"] / len(work) * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"serving: {len(work)} seeded 2D acoustic/elastic modeling requests, "
                         "NZ 100-600, NX 200-1500, NT 1000-5000, FD 4/8, Poisson arrivals"),
            "max_slots": args.slots, "cache_budget_mb": 192,
            "capacity_plain": cap_plain, "capacity_batched": cap_batch,
            "speedup_capacity": cap_batch["shots_per_hour"] / cap_plain["shots_per_hour"],
            "loads": loads,
            "with_gradients": mixed,
        },
    }
    print(json.dumps(out), flush=True)


def forward2d_main(args):
    """2D elastic forward modeling, one shot: 200 x 400, dh 10 m, dt 1 ms,
    NT 2000, FD4, CPML 20, vz Ricker 10 Hz, 400 receivers (the reference's
    CPU-runnable configuration).  A step is one time step; the whole shot is
    timed with the fields resident, and the oracle models the whole shot on
    the host cores."""
    from seiscl_amd import Context, Model
    from seiscl_amd.workloads import config1
    csts, params, shots = config1()
    NT, ncell = int(csts["NT"]), int(np.prod(csts["N"]))
    ctx = Context(0)
    m = Model(ctx, params, **csts)
    m.bind_shot(*shots[0])
    m.run_steps(0, NT)                       # warm-up: one whole shot
    m.bind_shot(*shots[0])
    ctx.sync()
    t0 = time.perf_counter()
    m.run_steps(0, NT)                       # replayed step graphs, no per-kernel events
    ctx.sync()
    wall = time.perf_counter() - t0
    # kernel times from a second shot with events around every update
    m.bind_shot(*shots[0])
    m.set_profiling(True)
    m.run_steps(0, NT)
    tm = m.timing()
    m.set_profiling(False)
    m.close()
    avg_v, avg_s = tm["update_v_ms"] / NT, tm["update_s_ms"] / NT
    ach = 44 * ncell / (avg_s / 1e3) / 1e9
    out = {
        "metric": METRIC, "value": round(ncell * NT / wall / 1e9, 4), "unit": "Gcell/s",
        "n_gpus": 1, "steps": NT, "warmup": NT, "ms_per_step": round(wall / NT * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "2D elastic forward modeling, one shot, 200x400, FD4, CPML 20, "
                               "NT 2000, 400 receivers", "grid": [200, 400], "NT": NT,
                   "shot_s": round(wall, 4), "update_v_us": round(avg_v * 1e3, 3),
                   "update_s_us": round(avg_s * 1e3, 3)},
        "roofline": {"bound": "hbm", "kernel": "update_s (2D elastic stress update)",
                     "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": None},
    }
    ctx.close()
    if not args.no_cpu:
        import oracle
        threads = min(16, os.cpu_count() or 1)
        t0 = time.time()
        oracle.forward(params, *shots[0], nthreads=threads, **csts)
        wall_c = time.time() - t0
        loop = oracle.last_loop_seconds()
        out["cpu_baseline"] = dict(
            value=ncell * NT / loop / 1e9, unit="Gcell/s", cores=threads, kind="port",
            sample=(f"the whole shot ({NT} steps, loop {loop:.2f} s, {wall_c:.2f} s with "
                    "set-up), C oracle with OpenMP"))
    print(json.dumps(out), flush=True)


BYTES_ADJ_S_2D = 108   # adjoint stress kernel, 2D elastic FP32, per cell
BYTES_ADJ_V_2D = 80    # adjoint velocity kernel, 2D elastic FP32, per cell


def gradient_main(args):
    """2D elastic FWI gradient (300 x 800, FD8, NT 3000, 16 shots) through the
    torch.autograd.Function: forward with boundary saving for every shot,
    backward = adjoint runs (and recomputed forward runs for all shots but
    the last).  A step is one whole gradient (forward + backward)."""
    import torch
    from seiscl_amd import Context, Model
    from seiscl_amd.autograd import Propagator, seismic
    from seiscl_amd.workloads import config3
    csts, true, start, shots = config3(NT=args.nt or 3000)
    if args.gradfix is not None:
        csts = dict(csts, gradfix=args.gradfix)
    NT, ncell = int(csts["NT"]), int(np.prod(csts["N"]))
    ctx = Context(0)
    t0 = time.time()
    mt = Model(ctx, true, **csts)
    din = [mt.forward(*sh) for sh in shots]
    mt.close()
    log(f"observed data of {len(shots)} shots in {time.time() - t0:.1f} s")
    obs = [torch.from_numpy(x) for d in din for x in d]
    prop = Propagator(ctx, start, **csts)
    params = [torch.tensor(start[n], requires_grad=True) for n in prop.param_names]

    def one():
        for p in params:
            p.grad = None
        ctx.sync()
        ta = time.perf_counter()
        outs = seismic(prop, shots, *params)
        tb = time.perf_counter()
        loss = sum(0.5 * ((o.double() - d.double()) ** 2).sum() for o, d in zip(outs, obs))
        tc = time.perf_counter()
        loss.backward()
        ctx.sync()
        td = time.perf_counter()
        return tb - ta, td - tc, (tc - tb), float(loss.detach())

    for _ in range(args.warmup):
        one()
    # timed without events (2D steps replay captured graphs), then one more
    # gradient with events around every adjoint kernel for their durations
    runs = [one() for _ in range(args.steps)]
    prop.model.set_profiling(True)
    one()
    tm = prop.model.timing()              # the last adjoint loop (shot 0)
    prop.model.set_profiling(False)
    grad = {n: p.grad.numpy().astype(np.float64) for n, p in zip(prop.param_names, params)}
    prop.close()
    fwd = sum(r[0] for r in runs) / len(runs)
    bwd = sum(r[1] for r in runs) / len(runs)
    misfit = sum(r[2] for r in runs) / len(runs)
    total = fwd + bwd + misfit
    cells = len(shots) * NT * ncell
    nadj = NT - 2
    adj_s, adj_v = tm["update_s_ms"] / nadj, tm["update_v_ms"] / nadj
    ach = BYTES_ADJ_S_2D * ncell / (adj_s / 1e3) / 1e9
    log(f"gradient: forward {fwd:.2f} s, backward {bwd:.2f} s, misfit {misfit:.2f} s; "
        f"adjoint kernels {adj_s * 1e3:.1f} / {adj_v * 1e3:.1f} us per step")
    cfg = {
        "workload": ("2D elastic FWI gradient through torch.autograd, 300x800, FD8, NT "
                     f"{NT}, 16 shots, CPML 16, boundary-saving recomputation"),
        "grid": [int(v) for v in csts["N"]], "fd_order": 8, "shots": len(shots), "NT": NT,
        "gradfix": int(csts.get("gradfix", 223)),
        "forward_s": round(fwd, 4), "backward_s": round(bwd, 4), "misfit_s": round(misfit, 4),
        "forward_gcell_s": round(cells / fwd / 1e9, 3),
        "backward_gcell_s": round(cells / bwd / 1e9, 3),
        "adj_s_us": round(adj_s * 1e3, 3), "adj_v_us": round(adj_v * 1e3, 3),
        "adj_v_gbs": round(BYTES_ADJ_V_2D * ncell / (adj_v / 1e3) / 1e9, 1),
        "misfit": runs[-1][3],
    }
    if args.fd_cells > 0:
        cfg["finite_differences"] = fd_check(ctx, csts, start, shots, din, grad, args.fd_cells)
    out = {
        "metric": METRIC, "value": round(cells / total / 1e9, 4), "unit": "Gcell/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(total * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": cfg,
        "roofline": {"bound": "hbm", "kernel": "adjoint stress update (2D elastic, FD8)",
                     "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": None},
    }
    ctx.close()
    if not args.no_cpu:
        import oracle
        threads = min(16, os.cpu_count() or 1)
        sp, s, rp = shots[0]
        nt = min(NT, args.cpu_steps)
        c = dict(csts, NT=nt)
        t0 = time.time()
        oracle.gradient(start, sp, s[:nt], rp, din=[d[:nt] for d in din[0]], nthreads=threads,
                        **c)
        wall = time.time() - t0
        out["cpu_baseline"] = dict(
            value=ncell * nt / wall / 1e9, unit="Gcell/s", cores=threads, kind="port",
            sample=(f"one shot's gradient (forward with boundary saving + adjoint), {nt} steps "
                    f"of the same grid in {wall:.1f} s, C oracle with OpenMP"))
    print(json.dumps(out), flush=True)


# viscoelastic adjoint kernels, 3D L=2 FD4, per cell: adjoint velocities read,
# adjoint stresses and memory variables read-modify-write, 10 coefficient
# arrays, the forward stresses of the step (FP16: 2 B, FP32: 4 B each), the
# gradient moduli, M and mu accumulators (double) read-modify-write
BYTES_ADJH_S_3D = {2: 248, 0: 260, 1: 260}
BYTES_ADJH_V_3D = {2: 82, 0: 88, 1: 88}


def gradient3d_main(args):
    """3D viscoelastic FWI gradient with FP16 wavefield storage: n^3 (512),
    NT 2500, L=2, FD4, CPML 20, free surface, 8 shots.  Per shot: forward run
    laying down level-1 checkpoints, then the adjoint (recomputed segments
    and adjoint steps); the adjoint sources are the shot's own seismograms
    (gradient of 1/2 |u|^2, the inputres mode), so no observed data need
    modeling.  Gradients accumulate in double over the shots."""
    from seiscl_amd import Context, Model
    from seiscl_amd.workloads import config5
    n = args.n or 512
    NT = args.nt or 2500
    t0 = time.time()
    csts, params, shots = config5(n=n, NT=NT, nshots=args.shots)
    log(f"workload n={n} NT={NT} shots={len(shots)} built in {time.time() - t0:.1f} s")
    ctx = Context(0)
    ncell = n ** 3

    def run(fp16, shots_run):
        t0 = time.time()
        m = Model(ctx, params, **dict(csts, FP16=fp16))
        log(f"FP16={fp16}: model resident in {time.time() - t0:.1f} s")
        m.set_profiling(True)
        m.grad_zero()
        rows = []
        for k, sh in enumerate(shots_run):
            ctx.sync()
            ta = time.perf_counter()
            out = m.forward_save(*sh)
            tb = time.perf_counter()
            m.adjoint(out)
            ctx.sync()
            tc = time.perf_counter()
            tm = m.timing()
            rows.append((tb - ta, tc - tb, tm))
            log(f"FP16={fp16} shot {k}: forward {tb - ta:.2f} s, adjoint {tc - tb:.2f} s, "
                f"plan {m.grad_plan()}, footprint {m.footprint() / 2**30:.1f} GiB")
        g = m.grad_read()
        plan, fp = m.grad_plan(), m.footprint()
        m.close()
        return rows, g, plan, fp

    rows, g16, plan, fp = run(args.fp16, shots)
    fwd = sum(r[0] for r in rows)
    adj = sum(r[1] for r in rows)
    nadj = NT - 1
    adj_s = sum(r[2]["update_s_ms"] for r in rows) / (len(rows) * nadj)
    adj_v = sum(r[2]["update_v_ms"] for r in rows) / (len(rows) * nadj)
    b_s, b_v = BYTES_ADJH_S_3D[args.fp16], BYTES_ADJH_V_3D[args.fp16]
    ach = b_s * ncell / (adj_s / 1e3) / 1e9
    cfg = {
        "workload": (f"3D viscoelastic FWI gradient, {n}^3, L=2, FD4, CPML 20, free surface, "
                     f"NT {NT}, {len(shots)} shots, FP16={args.fp16} storage, FP64 "
                     "accumulation, checkpoints + recomputation"),
        "grid": [n, n, n], "NT": NT, "shots": len(shots), "storage": args.fp16,
        "forward_s": round(fwd, 3), "adjoint_s": round(adj, 3),
        "forward_gcell_s": round(len(shots) * NT * ncell / fwd / 1e9, 3),
        "checkpoint_plan": plan, "footprint_gib": round(fp / 2**30, 2),
        "adj_s_ms": round(adj_s, 4), "adj_v_ms": round(adj_v, 4),
        "adj_v_gbs": round(b_v * ncell / (adj_v / 1e3) / 1e9, 1),
    }
    def rel(a, b):
        return {k: float(np.linalg.norm(a[k].astype(np.float64) - b[k])
                         / np.linalg.norm(b[k].astype(np.float64))) for k in b}

    if args.check_fp32 or args.check_fp64:
        # the first shot again in FP32 storage (and in FP64): gradient error
        # of the storage and of FP32 arithmetic, every parameter
        _, g32, p32, _ = run(0, shots[:1])
        _, g1, _, _ = run(args.fp16, shots[:1])
        cfg["fp32_plan"] = p32
        cfg["rel_l2_vs_fp32"] = rel(g1, g32)
        log(f"gradient error against FP32 storage (shot 0): {cfg['rel_l2_vs_fp32']}")
        if args.check_fp64:
            _, g64, p64, _ = run(64, shots[:1])
            cfg["fp64_plan"] = p64
            cfg["rel_l2_vs_fp64"] = {f"fp16={args.fp16}": rel(g1, g64), "fp32": rel(g32, g64)}
            log(f"gradient error against FP64 (shot 0): {cfg['rel_l2_vs_fp64']}")
    total = fwd + adj
    out = {
        "metric": METRIC, "value": round(len(shots) * NT * ncell / total / 1e9, 4),
        "unit": "Gcell/s", "n_gpus": 1, "steps": len(shots), "warmup": 0,
        "ms_per_step": round(total / len(shots) * 1e3, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": cfg,
        "roofline": {"bound": "hbm", "kernel": "viscoelastic adjoint stress update (3D, L=2)",
                     "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": None},
    }
    ctx.close()
    print(json.dumps(out), flush=True)


def fd_check(ctx, csts, start, shots, din, grad, ncells):
    """Central differences (h = 1e-3 of the value) of J = 1/2 sum (out - d)^2
    from FP64 runs, on ncells random interior cells per parameter."""
    from seiscl_amd import Model
    NT = int(csts["NT"])
    m64 = Model(ctx, start, **dict(csts, FP16=64))

    def J():
        s = 0.0
        for (sp, sr, rp), d in zip(shots, din):
            m64.bind_shot(sp, sr, rp)
            m64.run_steps(0, NT)
            s += 0.5 * sum(float(np.sum((o - x.astype(np.float64)) ** 2))
                           for o, x in zip(m64.read_shot(f64=True), d))
        return s

    rng = np.random.default_rng(300)
    nab = int(csts["nab"])
    res = {}
    for name in ("vp", "vs", "rho"):
        fd, gr = [], []
        # random interior cells the wavefield reaches (|g| >= 1% of its max)
        inner = np.zeros(grad[name].shape, bool)
        inner[nab + 10:-nab - 10, nab + 10:-nab - 10] = True
        a = np.abs(grad[name]) * inner
        cand = np.argwhere(a >= 1e-2 * a.max())
        for k in rng.choice(len(cand), size=ncells, replace=False):
            iz, ix = (int(v) for v in cand[k])
            v0 = float(start[name][iz, ix])
            vals = []
            for sgn in (1, -1):
                p = {k: v.copy() for k, v in start.items()}
                p[name][iz, ix] = np.float32(v0 * (1 + sgn * 1e-3))
                m64.set_params(p)
                vals.append((J(), float(p[name][iz, ix])))
            fd.append((vals[0][0] - vals[1][0]) / (vals[0][1] - vals[1][1]))
            gr.append(float(grad[name][iz, ix]))
        fd, gr = np.array(fd), np.array(gr)
        res[name] = {"cells": ncells,
                     "rel_l2": float(np.linalg.norm(gr - fd) / np.linalg.norm(fd)),
                     "median_ratio": float(np.median(gr / fd))}
        log(f"finite differences {name}: rel L2 {res[name]['rel_l2']:.3e} over {ncells} cells")
        log(f"finite differences {name}: {res[name]}")
    m64.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload",
                    choices=["forward3d", "forward2d", "serving", "gradient2d", "gradient3d"],
                    default="forward3d")
    ap.add_argument("--shots", type=int, default=8, help="gradient3d: shots")
    ap.add_argument("--fp16", type=int, default=2,
                    help="gradient3d: wavefield storage (2 FP16, 1 FP32 scaled, 0 FP32)")
    ap.add_argument("--check-fp32", action="store_true",
                    help="gradient3d: rerun shot 0 with FP32 storage and report the error")
    ap.add_argument("--check-fp64", action="store_true",
                    help="gradient3d: also rerun shot 0 in FP64 and report both errors")
    ap.add_argument("--nt", type=int, default=0, help="gradient2d: time steps (default 3000)")
    ap.add_argument("--gradfix", type=int, default=None,
                    help="gradient2d: seis_csts.gradfix (default 223; 255 adds the chain "
                         "through the staggered averages)")
    ap.add_argument("--fd-cells", type=int, default=0,
                    help="gradient2d: finite-difference check on this many cells per parameter")
    ap.add_argument("--requests", type=int, default=256, help="serving: number of requests")
    ap.add_argument("--slots", type=int, default=48, help="serving: batch slots")
    ap.add_argument("--grad-share", type=float, default=0.125,
                    help="serving: share of the elastic requests sent as gradient requests "
                         "in the mixed measurement (0: none)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="forward3d: time steps (1500); gradient2d: whole gradients (2)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="forward3d: time steps (10); gradient2d: whole gradients (1)")
    ap.add_argument("--n", type=int, default=0,
                    help="grid size per axis (forward3d 256, gradient3d 512)")
    ap.add_argument("--cpu-steps", type=int, default=150)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the FP16-storage and FP64 runs of the same shot")
    ap.add_argument("--no-serving", action="store_true",
                    help="forward3d: skip the bounded serving measurement")
    ap.add_argument("--no-traffic", action="store_true",
                    help="forward3d: skip the rocprofv3 HBM counter passes")
    ap.add_argument("--traffic-probe", choices=["visco", "elastic"], default=None,
                    help=argparse.SUPPRESS)
    ap.add_argument("--calib-copy", action="store_true",
                    help="run a 512 MiB device-to-device copy first (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the whole grid of this workload lives on "
              f"one MI355X (requested --gpus {args.gpus}); multi-GPU runs are not part "
              "of this build.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "serving":
        return serving_main(args)
    grad = args.workload == "gradient2d"
    if args.steps is None:
        args.steps = 2 if grad else 1500
    if args.warmup is None:
        args.warmup = 1 if grad else 10
    if grad:
        return gradient_main(args)
    if args.workload == "gradient3d":
        return gradient3d_main(args)
    if args.workload == "forward2d":
        return forward2d_main(args)
    args.n = args.n or 256
    if args.traffic_probe:
        return traffic_probe(args)
    forward3d_main(args)


def gradient_mix(ctx, work, share):
    """Observed seismograms for a share of the elastic requests (a model 2%
    faster in vp), which makes them gradient requests; untimed set-up."""
    from seiscl_amd import Model
    observed = [None] * len(work)
    el = [i for i, (c, _, _) in enumerate(work) if int(c["ND"]) == 2]
    step = max(1, int(round(1.0 / share))) if share > 0 else 0
    for i in (el[::step] if step else []):
        csts, params, shot = work[i]
        true = {k: (v * 1.02 if k == "vp" else v) for k, v in params.items()}
        m = Model(ctx, true, **csts)
        observed[i] = m.forward(*shot)
        m.close()
    return observed


def serving_mixed(ctx, work, share, slots):
    """Capacity with gradient requests in the mix: the plain path runs every
    request alone in arrival order; the batched path batches the modeling
    requests and runs the gradient requests beside the batch."""
    from seiscl_amd.serving import make_requests, serve_batched, serve_plain, summarize
    obs = gradient_mix(ctx, work, share)
    ngrad = sum(o is not None for o in obs)
    p = summarize(serve_plain(ctx, make_requests(work, observed=obs)))
    b = summarize(serve_batched(ctx, make_requests(work, observed=obs), max_slots=slots))
    keep = ("shots_per_hour", "p50_latency_s", "p99_latency_s", "span_s")
    pick = lambda d: {k: round(d[k], 4) for k in keep if k in d}   # noqa: E731
    log(f"mixed ({ngrad} gradient requests of {len(work)}): plain {p['shots_per_hour']:.0f}/h, "
        f"batched {b['shots_per_hour']:.0f}/h")
    return {"gradient_requests": ngrad, "requests": len(work), "capacity_plain": pick(p),
            "capacity_batched": pick(b),
            "speedup_capacity": round(b["shots_per_hour"] / p["shots_per_hour"], 3)}


def serving_sample(ctx, nreq=64, slots=48):
    """Bounded serving measurement for the default line: capacity of the
    plain and batched paths (all requests queued at t=0) and a Poisson load
    at the plain path's capacity."""
    from seiscl_amd.serving import (make_requests, poisson_arrivals, serve_batched,
                                    serve_plain, summarize)
    from seiscl_amd.workloads import config4
    work = config4(nreq)
    serve_plain(ctx, make_requests(work[:4]))
    cap_p = summarize(serve_plain(ctx, make_requests(work)))
    cap_b = summarize(serve_batched(ctx, make_requests(work), max_slots=slots))
    arr = poisson_arrivals(len(work), cap_p["shots_per_hour"] / 3600.0)
    ld_p = summarize(serve_plain(ctx, make_requests(work, arr)))
    ld_b = summarize(serve_batched(ctx, make_requests(work, arr), max_slots=slots))
    keep = ("shots_per_hour", "p50_latency_s", "p99_latency_s", "span_s")
    pick = lambda d: {k: round(d[k], 4) for k in keep if k in d}   # noqa: E731
    log(f"serving ({nreq} requests): capacity plain {cap_p['shots_per_hour']:.0f}/h, batched "
        f"{cap_b['shots_per_hour']:.0f}/h; 1x load p99 plain {ld_p['p99_latency_s']:.2f} s, "
        f"batched {ld_b['p99_latency_s']:.2f} s")
    return {
        "workload": (f"{nreq} seeded 2D acoustic/elastic modeling requests, NZ 100-600, "
                     "NX 200-1500, NT 1000-5000, FD 4/8, FP32"),
        "capacity_plain": pick(cap_p), "capacity_batched": pick(cap_b),
        "speedup_capacity": round(cap_b["shots_per_hour"] / cap_p["shots_per_hour"], 3),
        "load_1x_plain": pick(ld_p), "load_1x_batched": pick(ld_b),
        "with_gradients": serving_mixed(ctx, work, 0.125, slots),
    }


def forward3d_main(args):
    from seiscl_amd.workloads import config2
    W, K = args.warmup, args.steps
    NT = max(1500, W + K)
    traffic, traffic_el, tnote = None, None, "skipped (--no-traffic)"
    if not args.no_traffic:
        # before this process touches the GPU: the probes are child processes
        traffic, tnote = measure_traffic(args.n)
        if traffic is not None:
            traffic_el, _ = measure_traffic(args.n, elastic=True)
        log(f"traffic: {traffic} {traffic_el} ({tnote})")

    from seiscl_amd import Context, Model
    if args.calib_copy:
        import torch
        x = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log("calibration copy: 3 x 512 MiB read + 512 MiB write")

    log(f"building workload n={args.n} NT={NT}")
    csts, params, shots = config2(n=args.n, NT=NT)
    ctx = Context(0)
    ncell = int(np.prod(csts["N"]))

    def run(fp16, c=csts, p=params, full=True):
        t0 = time.time()
        m = Model(ctx, p, **dict(c, FP16=fp16))
        log(f"FP16={fp16} L={c['L']}: model resident in {time.time() - t0:.1f} s, "
            f"{m.footprint() / 2**30:.2f} GiB")
        m.bind_shot(*shots[0])
        m.run_steps(0, W)                       # warm-up, untimed
        m.set_profiling(True)
        ctx.sync()
        t0 = time.perf_counter()
        m.run_steps(W, K)                       # returns after the stream drained
        ctx.sync()
        wall = time.perf_counter() - t0
        tm = m.timing()
        m.set_profiling(False)
        seis = None
        if full and W + K < NT:
            m.run_steps(W + K, NT - W - K)      # the rest of the shot, untimed
        if full:
            seis = m.read_shot()
            if not all(np.isfinite(x).all() for x in seis):
                raise RuntimeError(f"non-finite seismograms (FP16={fp16})")
        m.close()
        avg_s, avg_v = tm["update_s_ms"] / K, tm["update_v_ms"] / K
        log(f"FP16={fp16} L={c['L']}: timed {K} steps: wall {wall:.3f} s, loop(events) "
            f"{tm['loop_ms'] / 1e3:.3f} s, update_s {avg_s:.3f} ms, update_v {avg_v:.3f} ms")
        return dict(wall=wall, avg_s=avg_s, avg_v=avg_v, seis=seis,
                    gcell=ncell * K / wall / 1e9)

    main_run = run(0, full=not args.no_variants)
    wall, avg_s, avg_v = main_run["wall"], main_run["avg_s"], main_run["avg_v"]
    value = main_run["gcell"]
    achieved = BYTES_PER_CELL_S * ncell / (avg_s / 1e3) / 1e9

    # the metric's 3D elastic case: the same shot and model without attenuation
    ce, pe, _ = config2(n=args.n, NT=NT, elastic=True)
    el = run(0, ce, pe, full=False)
    el_s = BYTES_PER_CELL_S_EL * ncell / (el["avg_s"] / 1e3) / 1e9
    el_v = BYTES_PER_CELL_V * ncell / (el["avg_v"] / 1e3) / 1e9
    el_step = BYTES_PER_CELL_EL * ncell / (el["wall"] / K) / 1e9
    elastic = {
        "workload": "the same shot without attenuation (L=0): 3D elastic, 256^3, FD4",
        "gcell_s": round(el["gcell"], 4), "ms_per_step": round(el["wall"] / K * 1e3, 5),
        "update_s_ms": round(el["avg_s"], 5), "update_v_ms": round(el["avg_v"], 5),
        "update_s_frac": round(el_s / PEAK_HBM_GBS, 4),
        "update_v_frac": round(el_v / PEAK_HBM_GBS, 4),
        "step_frac": round(el_step / PEAK_HBM_GBS, 4),
        "bytes_per_cell": BYTES_PER_CELL_EL,
        "traffic": traffic_el,
    }
    log(f"elastic: {elastic['gcell_s']} Gcell/s, fractions update_s {elastic['update_s_frac']}, "
        f"update_v {elastic['update_v_frac']}, step {elastic['step_frac']}")
    variants = {}
    if not args.no_variants:
        # the same shot with FP16 wavefield storage and in FP64; seismogram
        # error (relative L2 over all traces of the whole shot) against FP64
        r16, r64 = run(2), run(64)

        def err(a, b):
            a = np.concatenate([x.ravel() for x in a]).astype(np.float64)
            b = np.concatenate([x.ravel() for x in b]).astype(np.float64)
            return float(np.linalg.norm(a - b) / np.linalg.norm(b))

        variants = {
            "accuracy_window": f"all {NT} steps of the shot",
            "fp32_rel_l2_vs_fp64": err(main_run["seis"], r64["seis"]),
            "fp16_storage": {"gcell_s": round(r16["gcell"], 4),
                             "ms_per_step": round(r16["wall"] / K * 1e3, 5),
                             "update_s_ms": round(r16["avg_s"], 5),
                             "update_v_ms": round(r16["avg_v"], 5),
                             "bytes_per_cell": BYTES_PER_CELL_FP16,
                             "rel_l2_vs_fp64": err(r16["seis"], r64["seis"])},
            "fp64": {"gcell_s": round(r64["gcell"], 4),
                     "ms_per_step": round(r64["wall"] / K * 1e3, 5)},
        }
        log(f"accuracy vs FP64 over {NT} steps: FP32 {variants['fp32_rel_l2_vs_fp64']:.2e}, "
            f"FP16 storage {variants['fp16_storage']['rel_l2_vs_fp64']:.2e}")
    serving = None
    if not args.no_serving:
        serving = serving_sample(ctx)
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(wall / K * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"3D viscoelastic forward modeling, one shot, {args.n}^3 grid, "
                         "L=2, FD order 4, CPML 16, free surface, 64x64 receivers"),
            "grid": list(csts["N"].tolist()),
            "fd_order": 4, "L": 2, "precision": "fp32", "NT": NT,
            "update_v_ms": round(avg_v, 5), "update_s_ms": round(avg_s, 5),
            "update_v_gbs": round(BYTES_PER_CELL_V * ncell / (avg_v / 1e3) / 1e9, 1),
            "traffic_update_v": (traffic or {}).get("update_v"),
            "traffic_method": tnote,
            "elastic": elastic,
            **variants,
            "serving": serving,
        },
        "roofline": {
            "bound": "hbm", "kernel": "update_s (3D viscoelastic stress update)",
            "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": (traffic or {}).get("update_s"),
        },
    }
    ctx.close()
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {args.cpu_steps} steps on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(csts, params, shots[0], args.cpu_steps, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
