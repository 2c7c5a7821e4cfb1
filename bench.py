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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256, help="grid size per axis")
    ap.add_argument("--cpu-steps", type=int, default=150)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run a 512 MiB device-to-device copy first (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the whole grid of this workload lives on "
              f"one MI355X (requested --gpus {args.gpus}); multi-GPU runs are not part "
              "of this build.", file=sys.stderr)
        sys.exit(2)

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
    t0 = time.time()
    m = Model(ctx, params, **csts)
    log(f"model resident in {time.time() - t0:.1f} s, {m.footprint() / 2**30:.2f} GiB")
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
    if not all(np.isfinite(x).all() for x in seis):
        raise RuntimeError("non-finite seismograms")
    ncell = int(np.prod(csts["N"]))
    value = ncell * K / wall / 1e9
    avg_s = tm["update_s_ms"] / K
    avg_v = tm["update_v_ms"] / K
    achieved = BYTES_PER_CELL_S * ncell / (avg_s / 1e3) / 1e9
    log(f"timed {K} steps: wall {wall:.3f} s, loop(events) {tm['loop_ms'] / 1e3:.3f} s, "
        f"update_s {avg_s:.3f} ms, update_v {avg_v:.3f} ms per step")
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
        },
        "roofline": {
            "bound": "hbm", "kernel": "update_s (3D viscoelastic stress update)",
            "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": measured_traffic(args.n),
        },
    }
    m.close()
    ctx.close()
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {args.cpu_steps} steps on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(csts, params, shots[0], args.cpu_steps, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
