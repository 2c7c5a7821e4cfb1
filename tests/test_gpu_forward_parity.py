"""GPU forward modeling against the CPU oracle on the same seeded inputs.

Both sides evaluate every expression in the reference's order with FMA
contraction disabled, so seismograms and final wavefields are expected to
match bit for bit; the acceptance bar written here is relative L2 <= 1e-6
(the FP32 target is 1e-5), and a separate test checks bitwise equality.
"""
import numpy as np
import pytest

from cases import rel_l2, small_case

pytestmark = pytest.mark.gpu

TOL = 1e-6

CASES = (
    # 2D elastic, every FD order, CPML and Cerjan, with and without free surface
    [dict(ND=2, FDORDER=o, abs_type=a, freesurf=f) for o in (2, 4, 6, 8, 10, 12)
     for a in (1, 2) for f in (0, 1)]
    # 2D viscoelastic
    + [dict(ND=2, FDORDER=o, abs_type=1, freesurf=f, L=l) for o in (4, 8) for f in (0, 1)
       for l in (1, 2, 3)]
    # 2D acoustic
    + [dict(ND=22, FDORDER=o, abs_type=a, L=l) for o in (4, 8) for a in (0, 1, 2) for l in (0, 1)]
    # 3D elastic and viscoelastic
    + [dict(ND=3, FDORDER=o, abs_type=a, freesurf=f, L=l) for o in (2, 4, 8, 12)
       for a in (1, 2) for f in (0, 1) for l in (0, 2)]
    # every seismogram output set
    + [dict(ND=2, FDORDER=4, seisout=s) for s in (2, 3, 4)]
    + [dict(ND=3, FDORDER=4, seisout=s, L=1) for s in (2, 3, 4)]
    + [dict(ND=22, FDORDER=4, seisout=s) for s in (1, 4)]
    # Taylor coefficients, force source in x
    + [dict(ND=2, FDORDER=8, MAXRELERROR=0, src_type=0), dict(ND=3, FDORDER=6, MAXRELERROR=3)]
    # 3D grids several z tiles (of 64 cells) tall, some tiles clear of the
    # absorbing layers, and no absorbing layer at all
    + [dict(ND=3, FDORDER=o, abs_type=a, freesurf=f, L=l, N=(150, 28, 30))
       for o in (4, 8) for a in (1, 2) for f in (0, 1) for l in (0, 2)]
    + [dict(ND=3, FDORDER=4, abs_type=0, freesurf=f, N=(150, 28, 30)) for f in (0, 1)]
    # z extents that are multiples of 64: every wave full, z-derivatives
    # taken from wave neighbours
    + [dict(ND=3, FDORDER=o, abs_type=a, freesurf=f, L=l, N=(128, 20, 22))
       for o in (2, 4, 8, 12) for a in (1, 2) for f in (0, 1) for l in (0, 2)]
    + [dict(ND=3, FDORDER=4, abs_type=0, freesurf=f, N=(64, 20, 22)) for f in (0, 1)]
)


def _ids(c):
    return "-".join(f"{k}{v}" for k, v in c.items())


def _run_both(oracle_lib, gpu_ctx, case):
    from seiscl_amd import Model
    csts, params, (sp, s, rp) = small_case(**case)
    print(f"[parity] case={_ids(case)} N={tuple(csts['N'])}", flush=True)
    ref, fields = oracle_lib.forward(params, sp, s, rp, fields=True, **csts)
    m = Model(gpu_ctx, params, **csts)
    out = m.forward(sp, s, rp)
    gf = {k: m.read_field(k) for k in fields}
    m.close()
    return ref, out, fields, gf


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_forward_matches_oracle(oracle_lib, gpu_ctx, case):
    ref, out, fields, gf = _run_both(oracle_lib, gpu_ctx, case)
    assert len(ref) == len(out)
    for r, o in zip(ref, out):
        assert np.abs(r).max() > 0
        assert rel_l2(o, r) <= TOL
    for k in fields:
        assert rel_l2(gf[k], fields[k]) <= TOL, k


@pytest.mark.parametrize("case", [CASES[1], CASES[26], CASES[81], CASES[77], CASES[-9]], ids=_ids)
def test_forward_bitwise(oracle_lib, gpu_ctx, case):
    ref, out, fields, gf = _run_both(oracle_lib, gpu_ctx, case)
    for r, o in zip(ref, out):
        np.testing.assert_array_equal(o, r)
    for k in fields:
        np.testing.assert_array_equal(gf[k], fields[k])


def test_config1_full_size_bitwise(oracle_lib, gpu_ctx):
    """The 2D elastic configuration at its full size (200 x 400, dh 10 m,
    dt 1 ms, NT 2000, FD4, CPML 20, vz Ricker 10 Hz at (50 m, 2000 m), 400
    receivers at z = 50 m, seeded +-5% smooth perturbations): seismograms
    equal the oracle's bit for bit."""
    from seiscl_amd import Model
    from seiscl_amd.workloads import config1
    csts, params, shots = config1()
    print("[forward parity] config1 full size", flush=True)
    ref = oracle_lib.forward(params, *shots[0], nthreads=16, **csts)
    m = Model(gpu_ctx, params, **csts)
    out = m.forward(*shots[0])
    m.close()
    assert len(out) == 2
    for o, r in zip(out, ref):
        assert np.abs(r).max() > 0
        np.testing.assert_array_equal(o, r)


@pytest.mark.parametrize("case", [dict(ND=2, FDORDER=4, freesurf=1, L=1, NT=300),
                                  dict(ND=22, FDORDER=8, NT=200),
                                  dict(ND=2, FDORDER=4, seisout=4, NT=130)], ids=_ids)
def test_step_graph_replays_equal_eager(oracle_lib, gpu_ctx, case):
    """2D shots replay captured graphs of 64 steps (profiling off); with
    events around every kernel (profiling on) the same steps are launched
    one by one.  Several runs of each kind, interleaved on the same model
    with host work and synchronisations between them, give the same
    seismograms bit for bit, and those equal the oracle's; a run of steps
    started mid-shot (run_steps(t0, n)) replays from that step."""
    from seiscl_amd import Model
    csts, params, (sp, s, rp) = small_case(**case)
    ref = oracle_lib.forward(params, sp, s, rp, **csts)
    m = Model(gpu_ctx, params, **csts)
    outs = []
    for prof in (False, True, False, False, True, False):
        m.set_profiling(prof)
        outs.append(m.forward(sp, s, rp))
        gpu_ctx.sync()
        assert all(np.isfinite(o).all() for o in outs[-1])
    # the same shot advanced in uneven pieces
    m.set_profiling(False)
    NT = int(csts["NT"])
    m.bind_shot(sp, s, rp)
    t = 0
    for n in (5, 70, 64, NT):
        n = min(n, NT - t)
        if n <= 0:
            break
        m.run_steps(t, n)
        t += n
    outs.append(m.read_shot())
    m.close()
    for o in outs:
        for a, b in zip(o, outs[0]):
            np.testing.assert_array_equal(a, b)
    for a, r in zip(outs[0], ref):
        assert rel_l2(a, r) <= TOL
