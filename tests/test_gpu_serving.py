"""Batched serving on the GPU: every job's result is identical whether it
runs alone (the plain per-shot path, itself pinned against the oracle by
the forward parity tests) or inside a batch with other jobs of different
sizes, lengths, physics and FD orders; jobs join and leave the batch while
others are running."""
import numpy as np
import pytest

from cases import rel_l2

pytestmark = pytest.mark.gpu


def _requests(n, seed=9):
    from seiscl_amd.workloads import config4_request
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n):
        NZ = int(rng.choice([60, 84, 100]))
        NX = int(rng.choice([70, 96, 130]))
        NT = int(rng.choice([90, 140, 200]))
        reqs.append(config4_request(rng, nab=10, NZ=NZ, NX=NX, NT=NT))
    return reqs


def test_batch_equals_alone(oracle_lib, gpu_ctx):
    from seiscl_amd import Batch, Model
    reqs = _requests(6)
    kinds = {(c["ND"], c["FDORDER"]) for c, _, _ in reqs}
    print(f"[serving] kinds={sorted(kinds)}", flush=True)
    alone = []
    for csts, params, shot in reqs:
        m = Model(gpu_ctx, params, **csts)
        alone.append(m.forward(*shot))
        m.close()
    # one request against the oracle as well
    ref = oracle_lib.forward(reqs[0][1], *reqs[0][2], **reqs[0][0])
    for r, o in zip(ref, alone[0]):
        assert rel_l2(o, r) <= 1e-6
    batch = Batch(gpu_ctx, 8)
    models, slots = [], []
    for csts, params, shot in reqs:
        m = Model(gpu_ctx, params, **csts)
        m.bind_shot(*shot)
        slots.append(batch.add(m))
        models.append(m)
    while batch.status()[0]:
        batch.step(max(1, min(batch.status()[1], 37)))
    for m, a in zip(models, alone):
        for x, y in zip(m.read_shot(), a):
            np.testing.assert_array_equal(x, y)
    for s, m in zip(slots, models):
        batch.remove(s)
        m.close()
    batch.close()


def test_continuous_batching_staggered_arrivals(gpu_ctx):
    from seiscl_amd import Model
    from seiscl_amd.serving import make_requests, serve_batched, summarize
    work = _requests(10, seed=3)
    alone = []
    for csts, params, shot in work:
        m = Model(gpu_ctx, params, **csts)
        alone.append(m.forward(*shot))
        m.close()
    arrivals = np.linspace(0.0, 0.05, len(work))
    reqs = serve_batched(gpu_ctx, make_requests(work, arrivals), max_slots=4, chunk=25)
    for r, a in zip(sorted(reqs, key=lambda q: q.rid), alone):
        assert r.finish >= r.arrival
        for x, y in zip(r.out, a):
            np.testing.assert_array_equal(x, y)
    s = summarize(reqs)
    assert s["shots_per_hour"] > 0 and s["p99_latency_s"] >= s["p50_latency_s"]


def test_config4_size_requests(oracle_lib, gpu_ctx):
    """Requests at the serving workload's own sizes (NZ 400-600, NX
    1000-1500, NT 1000-1200, acoustic and elastic, FD 4 and 8, CPML 16):
    each large request equals the oracle, and every request of the mix --
    the large ones and eight drawn like the workload -- gives bit for bit
    the same seismograms in the batch as alone."""
    from seiscl_amd import Batch, Model
    from seiscl_amd.workloads import config4_request
    rng = np.random.default_rng(40)    # the large ones: elastic FD8, acoustic FD4
    big = [config4_request(rng, NZ=400, NX=1000, NT=1000),
           config4_request(rng, NZ=600, NX=1500, NT=1200)]
    # the workload's draw, with NT cut to 1500 to bound the run
    mix = [config4_request(rng, NT=int(rng.integers(1000, 1501))) for _ in range(8)]
    reqs = big + mix
    print(f"[serving] config-4 sizes: {[tuple(c['N']) + (c['NT'], c['ND']) for c, _, _ in reqs]}",
          flush=True)
    alone = []
    for csts, params, shot in reqs:
        m = Model(gpu_ctx, params, **csts)
        alone.append(m.forward(*shot))
        m.close()
    for (csts, params, shot), a in zip(big, alone):
        ref = oracle_lib.forward(params, *shot, nthreads=16, **csts)
        for r, o in zip(ref, a):
            assert np.abs(r).max() > 0
            assert rel_l2(o, r) <= 1e-6
    batch = Batch(gpu_ctx, len(reqs))
    models, slots = [], []
    for csts, params, shot in reqs:
        m = Model(gpu_ctx, params, **csts)
        m.bind_shot(*shot)
        slots.append(batch.add(m))
        models.append(m)
    while batch.status()[0]:
        batch.step(max(1, min(batch.status()[1], 64)))
    for m, a in zip(models, alone):
        for x, y in zip(m.read_shot(), a):
            np.testing.assert_array_equal(x, y)
    for s, m in zip(slots, models):
        batch.remove(s)
        m.close()
    batch.close()


def test_gradient_and_unbatched_requests(gpu_ctx):
    """Gradient requests, and modeling requests the batch kernels do not
    cover (viscoelastic, free surface), are served beside the batch on a
    second context of the GPU: each one's seismograms, gradient and misfit
    equal those of the same request run alone bit for bit, and the batched
    requests arriving among them still equal theirs."""
    from cases import small_case
    from seiscl_amd import Model
    from seiscl_amd.serving import Request, batchable, make_requests, run_alone, serve_batched
    work = _requests(6, seed=5)
    solo = [small_case(ND=2, FDORDER=4, L=0, NT=150, seed=1),
            small_case(ND=2, FDORDER=4, L=1, abs_type=1, freesurf=1, NT=150, seed=2),
            small_case(ND=2, FDORDER=8, L=2, freesurf=1, NT=120, seed=3)]
    observed = [None] * len(work)
    for k, (csts, params, shot) in enumerate(solo):
        work.append((csts, params, shot))
        if k == 2:
            observed.append(None)           # modeling outside the batch kernels
            continue
        true = {n: (v * 1.03 if n == "vp" else v) for n, v in params.items()}
        m = Model(gpu_ctx, true, **csts)
        observed.append(m.forward(*shot))
        m.close()
    assert not batchable(work[-1][0]) and batchable(work[0][0])
    alone = []
    for i, (csts, params, shot) in enumerate(work):
        r = Request(i, csts, params, shot, din=observed[i])
        run_alone(gpu_ctx, r)
        alone.append(r)
    arrivals = np.linspace(0.0, 0.04, len(work))
    order = np.random.default_rng(7).permutation(len(work))   # kinds interleaved in time
    reqs = serve_batched(gpu_ctx, make_requests(work, arrivals[order], observed), max_slots=4,
                         chunk=30)
    kinds = [r.kind for r in sorted(reqs, key=lambda q: q.rid)]
    print(f"[serving] kinds {kinds}", flush=True)
    assert kinds.count("gradient") == 2
    for r, a in zip(sorted(reqs, key=lambda q: q.rid), alone):
        assert r.finish >= r.arrival
        for x, y in zip(r.out, a.out):
            np.testing.assert_array_equal(x, y)
        assert r.misfit == a.misfit
        assert sorted(r.grad) == sorted(a.grad)
        for k in a.grad:
            assert np.abs(a.grad[k]).max() > 0, k
            np.testing.assert_array_equal(r.grad[k], a.grad[k])
