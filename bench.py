#!/usr/bin/env python3
"""Benchmark: 3D viscoelastic forward modeling on one MI355X.

Workload: one shot on a 256^3 grid (dh 10 m, dt 0.8 ms, FD order 4, CPML 16,
free surface, L=2 Maxwell bodies with Q=50 at 5-25 Hz, layered model with 3%
smooth perturbations, 12 Hz Ricker explosion, 64 x 64 surface receivers),
FP32.  A "step" is one time step of the finite-difference loop over the
whole grid: sources, velocity update, stress update, free surface and
receiver sampling.  W warm-up steps run untimed, then K steps are timed with
the wavefields resident in HBM.

value = grid cells x K / timed seconds  (Gcell-updates/s).
roofline: the stress update (the dominant kernel), algorithmic bytes per
launch = 196 B per cell (3 velocities read, 6 stresses and 12 memory
variables read-modify-write, 10 coefficient arrays read) / its average
duration from HIP events on the compute stream.
roofline.traffic: HBM bytes per stress-update launch from rocprofv3 counters
(FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, factors from a 512 MiB copy and fill).
cpu_baseline: the C oracle on the host cores, same grid, a bounded number of
steps.
config.fp16_storage / config.fp64: the same shot run with FP16 wavefield and
memory-variable storage (FP32 arithmetic) and in FP64; seismogram errors are
relative L2 norms against the FP64 run (skip with --no-variants).
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
# HBM bytes per stress-update launch at 256^3, from the FETCH_SIZE / WRITE_SIZE
# rocprofv3 passes summarised in this file (the command is written there)
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "update_s_v_visco3d_256_hbm_counters.md")


def measured_traffic(n):
    """HBM bytes per k_s3d launch from the committed counter summary (256^3 only)."""
    if n != 256 or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("traffic_bytes_per_launch k_s3d:"):
            return int(line.split(":")[1])
    return None


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
            "max_slots": args.slots,
            "capacity_plain": cap_plain, "capacity_batched": cap_batch,
            "speedup_capacity": cap_batch["shots_per_hour"] / cap_plain["shots_per_hour"],
            "loads": loads,
        },
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["forward3d", "serving"], default="forward3d")
    ap.add_argument("--requests", type=int, default=256, help="serving: number of requests")
    ap.add_argument("--slots", type=int, default=48, help="serving: batch slots")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256, help="grid size per axis")
    ap.add_argument("--cpu-steps", type=int, default=150)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the FP16-storage and FP64 runs of the same shot")
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

    from seiscl_amd import Context, Model
    from seiscl_amd.workloads import config2

    if args.calib_copy:
        import torch
        x = torch.ones(128 * 2**20, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log("calibration copy: 3 x 512 MiB read + 512 MiB write")

    W, K = args.warmup, args.steps
    log(f"building workload n={args.n} NT={W + K}")
    csts, params, shots = config2(n=args.n, NT=W + K)
    ctx = Context(0)
    ncell = int(np.prod(csts["N"]))

    def run(fp16):
        t0 = time.time()
        m = Model(ctx, params, **dict(csts, FP16=fp16))
        log(f"FP16={fp16}: model resident in {time.time() - t0:.1f} s, "
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
        seis = m.read_shot()
        m.close()
        if not all(np.isfinite(x).all() for x in seis):
            raise RuntimeError(f"non-finite seismograms (FP16={fp16})")
        avg_s, avg_v = tm["update_s_ms"] / K, tm["update_v_ms"] / K
        log(f"FP16={fp16}: timed {K} steps: wall {wall:.3f} s, loop(events) "
            f"{tm['loop_ms'] / 1e3:.3f} s, update_s {avg_s:.3f} ms, update_v {avg_v:.3f} ms")
        return dict(wall=wall, avg_s=avg_s, avg_v=avg_v, seis=seis,
                    gcell=ncell * K / wall / 1e9)

    main_run = run(0)
    wall, avg_s, avg_v = main_run["wall"], main_run["avg_s"], main_run["avg_v"]
    value = main_run["gcell"]
    achieved = BYTES_PER_CELL_S * ncell / (avg_s / 1e3) / 1e9
    variants = {}
    if not args.no_variants:
        # the same shot with FP16 wavefield storage and in FP64; seismogram
        # error (relative L2 over all traces) against the FP64 run
        r16, r64 = run(2), run(64)

        def err(a, b):
            a = np.concatenate([x.ravel() for x in a]).astype(np.float64)
            b = np.concatenate([x.ravel() for x in b]).astype(np.float64)
            return float(np.linalg.norm(a - b) / np.linalg.norm(b))

        variants = {
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
        log(f"accuracy vs FP64: FP32 {variants['fp32_rel_l2_vs_fp64']:.2e}, "
            f"FP16 storage {variants['fp16_storage']['rel_l2_vs_fp64']:.2e}")
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
            "fd_order": 4, "L": 2, "precision": "fp32",
            "update_v_ms": round(avg_v, 5), "update_s_ms": round(avg_s, 5),
            "update_v_gbs": round(BYTES_PER_CELL_V * ncell / (avg_v / 1e3) / 1e9, 1),
            **variants,
        },
        "roofline": {
            "bound": "hbm", "kernel": "update_s (3D viscoelastic stress update)",
            "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": measured_traffic(args.n),
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
