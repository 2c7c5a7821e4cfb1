"""GPU gradient against the CPU oracle on the same inputs: elastic runs by
backpropagation with boundary saving, viscoelastic runs from checkpoints
and recomputed segments (the oracle keeps the whole forward history);
viscoelastic runs also compare the tau_p and tau_s gradients.

The adjoint kernels follow the oracle's expressions with FMA contraction
disabled, and both accumulate the per-step gradient increments in double in
the same time order, so seismograms and gradients are expected to agree bit
for bit; the acceptance bar written here is relative L2 <= 1e-6 (the FP32
target is 1e-5), and one test checks bitwise equality.
"""
import numpy as np
import pytest

from cases import rel_l2, small_case

pytestmark = pytest.mark.gpu

TOL = 1e-6

CASES = (
    [dict(ND=2, FDORDER=o, abs_type=a) for o in (2, 4, 8, 12) for a in (1, 2)]
    + [dict(ND=2, FDORDER=4, gradfix=f) for f in (0, 1, 2, 4)]
    + [dict(ND=2, FDORDER=4, seisout=2), dict(ND=2, FDORDER=8, seisout=1, MAXRELERROR=0)]
    + [dict(ND=3, FDORDER=o, abs_type=a) for o in (2, 4, 8) for a in (1, 2)]
    + [dict(ND=3, FDORDER=4, gradfix=0), dict(ND=3, FDORDER=4, seisout=2)]
    # free surface: adjoint image method, 3D with and without the syy fix
    + [dict(ND=2, FDORDER=o, freesurf=1) for o in (4, 8)]
    + [dict(ND=2, FDORDER=4, freesurf=1, abs_type=2)]
    + [dict(ND=3, FDORDER=4, freesurf=1, gradfix=f) for f in (15, 7)]
    # tall 3D grid: several z tiles of 64 cells
    + [dict(ND=3, FDORDER=4, abs_type=a, N=(150, 28, 30)) for a in (1, 2)]
    + [dict(ND=3, FDORDER=4, abs_type=1, freesurf=f, N=(128, 20, 22)) for f in (0, 1)]
)
# viscoelastic: forward state from checkpoints and recomputed segments
VISCO = (
    [dict(ND=2, FDORDER=4, L=l, abs_type=a) for l in (1, 2, 3) for a in (1, 2)]
    + [dict(ND=2, FDORDER=8, L=2), dict(ND=2, FDORDER=4, L=2, freesurf=1)]
    + [dict(ND=2, FDORDER=4, L=2, gradfix=f) for f in (0, 15)]
    + [dict(ND=3, FDORDER=4, L=2, abs_type=a) for a in (1, 2)]
    + [dict(ND=3, FDORDER=4, L=1, freesurf=1), dict(ND=3, FDORDER=8, L=2, gradfix=15)]
    # the benchmark gradient's case: 3D, L=2, FD4, free surface
    + [dict(ND=3, FDORDER=4, L=2, freesurf=1), dict(ND=3, FDORDER=4, L=2, freesurf=1, gradfix=223)]
    + [dict(ND=3, FDORDER=4, L=2, freesurf=1, N=(128, 20, 22))]
)
CASES = CASES + VISCO


def _ids(c):
    return "-".join(f"{k}{v}" for k, v in c.items())


def _setup(case):
    case = dict(case)
    gradfix = case.pop("gradfix", 31)
    extra = {k: case.pop(k) for k in ("Hout", "restype") if k in case}
    nt = 160 if case.get("ND") == 3 else 220
    csts, params, shot = small_case(NT=nt, seed=11, **case)
    csts["gradfix"] = gradfix
    csts.update(extra)
    rng = np.random.default_rng(4)
    true = {k: (v * (1 + 0.04 * rng.standard_normal(v.shape))).astype(np.float32)
            for k, v in params.items()}
    return csts, params, shot, true


def _run_both(oracle_lib, gpu_ctx, case):
    from seiscl_amd import Model
    csts, params, (sp, s, rp), true = _setup(case)
    print(f"[grad parity] case={_ids(case)} N={tuple(csts['N'])}", flush=True)
    din = oracle_lib.forward(true, sp, s, rp, **csts)
    ref = oracle_lib.gradient(params, sp, s, rp, din=din, **csts)
    m = Model(gpu_ctx, params, **csts)
    m.grad_zero()
    out, rms = m.gradient(sp, s, rp, din=din)
    g = m.grad_read()
    m.close()
    return ref, out, rms, g


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_gradient_matches_oracle(oracle_lib, gpu_ctx, case):
    ref, out, rms, g = _run_both(oracle_lib, gpu_ctx, case)
    for r, o in zip(ref["out"], out):
        assert rel_l2(o, r) <= TOL
    assert abs(rms[0] - ref["rms"][0]) <= 1e-9 * abs(ref["rms"][0])
    for k in ref["grad"]:
        assert np.abs(ref["grad"][k]).max() > 0, k
        assert rel_l2(g[k], ref["grad"][k]) <= TOL, k


@pytest.mark.parametrize("case", [dict(ND=2, FDORDER=4, abs_type=1),
                                  dict(ND=2, FDORDER=4, freesurf=1, abs_type=2),
                                  dict(ND=3, FDORDER=4, abs_type=2, N=(150, 28, 30)),
                                  dict(ND=2, FDORDER=4, L=2, abs_type=1),
                                  dict(ND=3, FDORDER=4, L=1, freesurf=1),
                                  dict(ND=3, FDORDER=4, L=2, freesurf=1)], ids=_ids)
def test_gradient_bitwise(oracle_lib, gpu_ctx, case):
    ref, out, rms, g = _run_both(oracle_lib, gpu_ctx, case)
    for r, o in zip(ref["out"], out):
        np.testing.assert_array_equal(o, r)
    for k in ref["grad"]:
        np.testing.assert_array_equal(g[k], ref["grad"][k])


@pytest.mark.parametrize("ND", [2, 3])
def test_adjoint_sources_given(oracle_lib, gpu_ctx, ND):
    """inputres mode: the adjoint sources are given (then scaled), and the
    split entry points (forward_save, adjoint, grad_read) give the same
    gradient as the one-call path."""
    from seiscl_amd import Model
    csts, params, (sp, s, rp), _ = _setup(dict(ND=ND, FDORDER=4))
    rng = np.random.default_rng(7)
    res = [rng.standard_normal((csts["NT"], rp.shape[1])).astype(np.float32)
           for _ in range(3 if ND == 3 else 2)]
    ref = oracle_lib.gradient(params, sp, s, rp, res=res, **csts)
    m = Model(gpu_ctx, params, **csts)
    m.grad_zero()
    m.gradient(sp, s, rp, res=res)
    g1 = m.grad_read()
    m.grad_zero()
    m.forward_save(sp, s, rp)
    m.adjoint(res)
    g2 = m.grad_read()
    m.close()
    for k in ref["grad"]:
        assert rel_l2(g1[k], ref["grad"][k]) <= TOL, k
        np.testing.assert_array_equal(g1[k], g2[k])


# (case, tight): tight runs the viscoelastic case under a checkpoint plan
# with one more level
MOVIE = [(dict(ND=2, FDORDER=4, abs_type=1), False), (dict(ND=3, FDORDER=4, freesurf=1), False),
         (dict(ND=2, FDORDER=4, L=2, abs_type=1), False),
         (dict(ND=2, FDORDER=4, L=2, abs_type=1), True),
         (dict(ND=3, FDORDER=4, L=1, freesurf=1), False),
         (dict(ND=3, FDORDER=4, L=1, freesurf=1), True)]


@pytest.mark.parametrize("case,tight", MOVIE, ids=lambda v: _ids(v) if isinstance(v, dict) else
                         ("tight" if v else "default"))
def test_gradient_movie(oracle_lib, gpu_ctx, case, tight):
    """Movie of a gradient run (seis_adjoint_movie, src/time_stepping.c:
    926-929): every frame equals the oracle's bit for bit -- backpropagated
    forward fields (elastic) or the recomputed forward state (viscoelastic,
    also under a checkpoint plan with one more level) -- and the gradient
    equals the one of the run without movie."""
    from seiscl_amd import Model
    csts, params, (sp, s, rp), _ = _setup(case)
    mv = 40
    print(f"[grad movie] case={_ids(case)} tight={tight}", flush=True)
    rng = np.random.default_rng(8)
    nout = 3 if case["ND"] == 3 else 2
    res = [rng.standard_normal((csts["NT"], rp.shape[1])).astype(np.float32)
           for _ in range(nout)]
    ref = oracle_lib.gradient(params, sp, s, rp, res=res, movout=mv, **csts)
    m = Model(gpu_ctx, params, **csts)
    names = list(ref["movie"])
    m.grad_zero()
    m.forward_save(sp, s, rp)
    if tight:
        plan = m.grad_plan()
        m.set_grad_memory(plan["bytes"] - 1)
        m.forward_save(sp, s, rp)
        assert m.grad_plan()["levels"] > plan["levels"]
    mov = m.adjoint(res, movout=mv, movie_vars=names)
    g1 = m.grad_read()
    m.grad_zero()
    m.forward_save(sp, s, rp)
    m.adjoint(res)
    g2 = m.grad_read()
    m.close()
    nfr = csts["NT"] // mv
    for name in names:
        a = mov[name]
        b = np.moveaxis(ref["movie"][name], 0, -1)
        assert a.shape == tuple(csts["N"]) + (nfr,)
        assert np.abs(b).max() > 0, name
        np.testing.assert_array_equal(a, b, err_msg=name)
    for k in g1:
        np.testing.assert_array_equal(g1[k], g2[k])
        assert rel_l2(g1[k], ref["grad"][k]) <= TOL, k


def test_gradient_accumulates_over_shots(oracle_lib, gpu_ctx):
    from seiscl_amd import Model
    csts, params, (sp, s, rp), true = _setup(dict(ND=2, FDORDER=4))
    shots = [(sp, s, rp), (sp[:, :1].copy(), s[:, :1].copy(), rp)]
    shots[1][0][0, 0] += 40.0
    refs = []
    for shot in shots:
        din = oracle_lib.forward(true, *shot, **csts)
        refs.append((din, oracle_lib.gradient(params, *shot, din=din, **csts)["grad"]))
    m = Model(gpu_ctx, params, **csts)
    m.grad_zero()
    for shot, (din, _) in zip(shots, refs):
        m.gradient(*shot, din=din)
    g = m.grad_read()
    m.close()
    for k in g:
        tot = refs[0][1][k].astype(np.float64) + refs[1][1][k]
        assert rel_l2(g[k], tot) <= 1e-5, k


# ---- viscoelastic gradients with scaled or FP16 wavefield storage and the
# ---- checkpoint plans

VISCO_SCALED = [dict(ND=3, FDORDER=4, L=2, abs_type=1),
                dict(ND=3, FDORDER=4, L=2, abs_type=1, freesurf=1),
                dict(ND=3, FDORDER=4, L=1, abs_type=2, freesurf=1),
                dict(ND=2, FDORDER=4, L=2, abs_type=1),
                dict(ND=2, FDORDER=8, L=1, abs_type=1, freesurf=1)]


def _gpu_grad(gpu_ctx, case, FP16=0, mem=0):
    from seiscl_amd import Model
    csts, params, (sp, s, rp), true = _setup(case)
    print(f"[grad visco] case={_ids(case)} FP16={FP16} mem={mem}", flush=True)
    m = Model(gpu_ctx, true, **csts)
    din = m.forward(sp, s, rp)
    m.close()
    m = Model(gpu_ctx, params, **dict(csts, FP16=FP16))
    m.set_grad_memory(mem)
    m.grad_zero()
    out, _ = m.gradient(sp, s, rp, din=din)
    g = m.grad_read()
    plan = m.grad_plan()
    m.close()
    return out, g, plan


@pytest.mark.parametrize("case", VISCO_SCALED, ids=_ids)
def test_visco_gradient_scaled_units(gpu_ctx, case):
    """FP16=1 stores FP32 wavefields in the scaled units of FP16 storage
    (velocities times av, stresses times as, coefficients rescaled by
    q = av/as, adjoint velocities times q).  Every factor is a power of two,
    so the runs differ only where FP32 values ahead of the wavefront are
    subnormal in one scaling and not in the other: seismograms within 1e-6
    and gradients within 1e-5 relative L2 of the FP32 run (a wrong unit
    factor would be off by O(1))."""
    o32, g32, _ = _gpu_grad(gpu_ctx, case, 0)
    o1, g1, _ = _gpu_grad(gpu_ctx, case, 1)
    for a, b in zip(o1, o32):
        assert rel_l2(a, b) <= 1e-6
    for k in g32:
        assert np.abs(g32[k]).max() > 0, k
        err = rel_l2(g1[k], g32[k])
        print(f"[grad visco] scaled units {k}: rel L2 {err:.2e}", flush=True)
        assert err <= 1e-5, k


@pytest.mark.parametrize("case", VISCO_SCALED, ids=_ids)
def test_visco_gradient_fp16_storage(gpu_ctx, case):
    """FP16 storage of the forward wavefields and memory variables (FP32
    arithmetic, FP32 adjoint fields, double accumulation) against the FP32
    run.  On this small grid (about a third of the cells in absorbing layers
    or next to the surface, 160 steps) the FP16 errors measure 1.1-1.8e-2 for
    vp, vs, rho and 2.1-2.9e-2 for taup (MI355X), against 3-5e-4 at 256^3
    over 500 steps; the bars are 2e-2 and, for the tau gradients, 4e-2."""
    o32, g32, _ = _gpu_grad(gpu_ctx, case, 0)
    o16, g16, _ = _gpu_grad(gpu_ctx, case, 2)
    for a, b in zip(o16, o32):
        print(f"[grad visco] FP16 storage seismogram: rel L2 {rel_l2(a, b):.2e}", flush=True)
        assert rel_l2(a, b) <= 1e-2
    for k in g32:
        err = rel_l2(g16[k], g32[k])
        print(f"[grad visco] FP16 storage {k}: rel L2 {err:.2e}", flush=True)
        assert err <= (4e-2 if k in ("taup", "taus") else 2e-2), k


FP64_CASES = [dict(ND=2, FDORDER=4), dict(ND=3, FDORDER=4, freesurf=1),
              dict(ND=2, FDORDER=8, L=1, abs_type=1, freesurf=1),
              dict(ND=3, FDORDER=4, L=2, abs_type=1, freesurf=1)]


@pytest.mark.parametrize("case", FP64_CASES, ids=_ids)
def test_gradient_fp64_runs(gpu_ctx, case):
    """FP64 runs (FP16=64: wavefields, adjoint fields, boundary shells,
    checkpoints and histories in double, double arithmetic) against the FP32
    run of the same kernels, which equals the oracle bit for bit: the FP32
    gradient's error is its rounding, within 1e-3 relative L2 of the FP64
    one on these grids, and the two runs are not the same computation."""
    o32, g32, _ = _gpu_grad(gpu_ctx, case, 0)
    o64, g64, _ = _gpu_grad(gpu_ctx, case, 64)
    for a, b in zip(o32, o64):
        assert rel_l2(a, b) <= 1e-5
    diff = 0.0
    for k in g32:
        assert np.abs(g64[k]).max() > 0, k
        err = rel_l2(g32[k], g64[k])
        print(f"[grad fp64] {_ids(case)} {k}: FP32 vs FP64 rel L2 {err:.2e}", flush=True)
        assert err <= 1e-3, k
        diff = max(diff, err)
    assert diff > 0


@pytest.mark.parametrize("case", [dict(ND=2, FDORDER=4, L=2, abs_type=1),
                                  dict(ND=3, FDORDER=4, L=2, abs_type=1, freesurf=1)], ids=_ids)
def test_visco_checkpoint_plans(gpu_ctx, case):
    """The default plan is the one-level plan of least memory.  A budget one
    byte smaller leaves no one-level plan and a two-level one takes over
    (its segments recomputed once more to lay down a second level of
    checkpoints); one byte below that, three levels.  The recomputed states
    are the same, so the gradients equal the default plan's bit for bit."""
    _, g0, p0 = _gpu_grad(gpu_ctx, case)
    assert p0["levels"] == 1 and p0["bytes"] > 0, p0
    plans = [p0]
    for levels in (2, 3):
        _, g, p = _gpu_grad(gpu_ctx, case, mem=plans[-1]["bytes"] - 1)
        print(f"[grad visco] plan {p}", flush=True)
        assert p["levels"] == levels and p["bytes"] < plans[-1]["bytes"], p
        for k in g0:
            assert np.abs(g0[k]).max() > 0, k
            np.testing.assert_array_equal(g[k], g0[k])
        plans.append(p)


def test_visco_checkpoint_budget_error(gpu_ctx):
    from seiscl_amd import Model
    from seiscl_amd._lib import SeisError
    csts, params, (sp, s, rp), _ = _setup(dict(ND=2, FDORDER=4, L=1))
    m = Model(gpu_ctx, params, **csts)
    m.set_grad_memory(1024)
    with pytest.raises(SeisError, match="no checkpoint plan"):
        m.forward_save(sp, s, rp)
    m.close()


RTM_HESS = [dict(ND=2, FDORDER=4, Hout=1, gradfix=223),
            dict(ND=2, FDORDER=4, restype=1, Hout=1, gradfix=223),
            dict(ND=3, FDORDER=4, Hout=1, freesurf=1, gradfix=223),
            dict(ND=3, FDORDER=4, restype=1, gradfix=223),
            dict(ND=2, FDORDER=4, L=2, Hout=1, gradfix=223),
            dict(ND=3, FDORDER=4, L=1, restype=1, Hout=1, gradfix=223),
            dict(ND=2, FDORDER=4, Hout=1, gradfix=31)]


@pytest.mark.parametrize("case", RTM_HESS, ids=_ids)
def test_rtm_and_hessian_match_oracle(oracle_lib, gpu_ctx, case):
    """restype 1 (RTM image of M) and Hout (approximate Hessian) against
    the oracle: gradients and Hessians bit for bit."""
    from seiscl_amd import Model
    csts, params, (sp, s, rp), true = _setup(case)
    print(f"[grad rtm/hess] case={_ids(case)}", flush=True)
    din = oracle_lib.forward(true, sp, s, rp, **csts)
    ref = oracle_lib.gradient(params, sp, s, rp, din=din, **csts)
    m = Model(gpu_ctx, params, **csts)
    m.grad_zero()
    m.gradient(sp, s, rp, din=din)
    g = m.grad_read()
    h = m.hess_read() if csts.get("Hout") else None
    m.close()
    for k in ref["grad"]:
        np.testing.assert_array_equal(g[k], ref["grad"][k])
    assert np.abs(ref["grad"]["vp"]).max() > 0
    if h is not None:
        for k in ref["hess"]:
            np.testing.assert_array_equal(h[k], ref["hess"][k])
        assert np.abs(ref["hess"]["vp"]).max() > 0


@pytest.mark.parametrize("case", [dict(ND=2, FDORDER=4), dict(ND=3, FDORDER=4, freesurf=1),
                                  dict(ND=2, FDORDER=4, L=1, abs_type=1),
                                  dict(ND=3, FDORDER=4, L=2, abs_type=1)], ids=_ids)
def test_staggered_chain_conserves_totals(gpu_ctx, case):
    """gradfix bit 32 moves the shear and density terms from the cell of
    the same index to the cells each staggered node averages over, with
    weights that sum to one over a node's cells when the model is
    homogeneous: on a homogeneous starting model the totals of the vs and
    rho gradients equal those of the reference's attribution (to rounding)
    while the cell values differ, and the vp gradient, which enters no
    staggered average, is the same bit for bit."""
    from seiscl_amd import Model
    csts, params, (sp, s, rp), true = _setup(dict(case, gradfix=223))
    homog = {k: np.full_like(v, float(np.median(v))) for k, v in params.items()}
    m = Model(gpu_ctx, true, **csts)
    din = m.forward(sp, s, rp)
    m.close()
    g = {}
    for gf in (223, 255):
        print(f"[grad chain] case={_ids(case)} gradfix={gf}", flush=True)
        m = Model(gpu_ctx, homog, **dict(csts, gradfix=gf))
        m.grad_zero()
        m.gradient(sp, s, rp, din=din)
        g[gf] = m.grad_read()
        m.close()
    np.testing.assert_array_equal(g[255]["vp"], g[223]["vp"])
    for k in ("vs", "rho"):
        a, b = g[255][k].astype(np.float64), g[223][k].astype(np.float64)
        assert abs(a.sum() - b.sum()) <= 1e-5 * np.abs(b).sum(), k
        assert rel_l2(a, b) > 1e-3, k


@pytest.mark.parametrize("case", [dict(ND=2, FDORDER=4, abs_type=1),
                                  dict(ND=2, FDORDER=8, freesurf=1),
                                  dict(ND=2, FDORDER=4, abs_type=2, FP16=64),
                                  dict(ND=2, FDORDER=4, tmin=0.0305)], ids=_ids)
def test_gradient_graph_replays_equal_eager(oracle_lib, gpu_ctx, case):
    """2D elastic gradients replay captured graphs of 64 forward steps with
    boundary saving and of 64 adjoint steps (profiling off); with profiling
    on every kernel is launched one by one.  Interleaved runs of both kinds
    on the same model give the same seismograms and gradients bit for bit,
    and (FP32) those equal the oracle's."""
    from seiscl_amd import Model
    case = dict(case)
    fp16 = case.pop("FP16", 0)
    tmin = case.pop("tmin", 0)
    csts, params, (sp, s, rp), true = _setup(case)
    if tmin:
        csts["tmin"] = tmin
    print(f"[grad graph] case={_ids(case)} FP16={fp16} tmin={tmin}", flush=True)
    din = oracle_lib.forward(true, sp, s, rp, **csts)
    m = Model(gpu_ctx, params, **dict(csts, FP16=fp16))
    runs = []
    for prof in (False, True, False, True, False):
        m.set_profiling(prof)
        m.grad_zero()
        out, _ = m.gradient(sp, s, rp, din=din)
        runs.append((out, m.grad_read()))
        gpu_ctx.sync()
    m.close()
    for out, g in runs[1:]:
        for a, b in zip(out, runs[0][0]):
            np.testing.assert_array_equal(a, b)
        for k in g:
            np.testing.assert_array_equal(g[k], runs[0][1][k])
    if fp16 == 0:
        ref = oracle_lib.gradient(params, sp, s, rp, din=din, **csts)
        for k in ref["grad"]:
            assert rel_l2(runs[0][1][k], ref["grad"][k]) <= TOL, k
