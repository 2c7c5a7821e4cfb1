"""DFT-domain gradient (back_prop_type 2) on the GPU against the oracle's
restatement (tests/test_oracle_dft.py pins the oracle): the spectra are
accumulated in float with the same factor table and order, the scalar
products formed in float and the coefficients in double, so the gradients
agree to rounding of the double coefficient arithmetic (bar 1e-5 relative
L2; MI355X: at most 4e-8), the seismograms bit for bit and the rms at the
gradient frequencies to 1e-9."""
import numpy as np
import pytest

from cases import rel_l2, small_case

pytestmark = pytest.mark.gpu

CASES = [dict(ND=2, FDORDER=4), dict(ND=2, FDORDER=8, abs_type=2),
         dict(ND=2, FDORDER=4, freesurf=1), dict(ND=2, FDORDER=4, param_type=1),
         dict(ND=2, FDORDER=4, param_type=2, gradfix=0),
         dict(ND=3, FDORDER=4), dict(ND=3, FDORDER=4, freesurf=1),
         dict(ND=2, FDORDER=4, L=1), dict(ND=2, FDORDER=4, L=2, freesurf=1),
         dict(ND=2, FDORDER=4, L=1, param_type=2), dict(ND=3, FDORDER=4, L=2, freesurf=1)]
FREQS = np.array([9.0, 15.0, 24.0], np.float32)   # bins 1-5 of both records


def _ids(c):
    return "-".join(f"{k}{v}" for k, v in c.items())


def _setup(case):
    case = dict(case)
    gradfix = case.pop("gradfix", 223)
    pt = case.pop("param_type", 0)
    csts, params, shot = small_case(NT=160 if case.get("ND") == 3 else 220, seed=11, **case)
    csts["gradfix"] = gradfix
    csts["param_type"] = pt
    if pt == 1:
        params = dict(params, M=params["vp"] ** 2 * params["rho"],
                      mu=params["vs"] ** 2 * params["rho"])
        del params["vp"], params["vs"]
    elif pt == 2:
        params = dict(params, Ip=params["vp"] * params["rho"], Is=params["vs"] * params["rho"])
        del params["vp"], params["vs"]
    rng = np.random.default_rng(4)
    true = {k: (v * (1 + 0.04 * rng.standard_normal(v.shape))).astype(np.float32)
            for k, v in params.items()}
    return csts, params, shot, true


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_dft_gradient_matches_oracle(oracle_lib, gpu_ctx, case):
    from seiscl_amd import Model
    csts, params, (sp, s, rp), true = _setup(case)
    print(f"[dft] case={_ids(case)}", flush=True)
    din = oracle_lib.forward(true, sp, s, rp, **csts)
    ref = oracle_lib.gradient_dft(params, sp, s, rp, FREQS, din=din, **csts)
    m = Model(gpu_ctx, params, **csts)
    m.grad_dft(FREQS)
    m.grad_zero()
    out, rms = m.gradient(sp, s, rp, din=din)
    g = m.grad_read()
    m.close()
    for r, o in zip(ref["out"], out):
        np.testing.assert_array_equal(o, r)
    assert rms[0] == pytest.approx(ref["rms"][0], rel=1e-9)
    assert rms[1] == pytest.approx(ref["rms"][1], rel=1e-9)
    for k in ref["grad"]:
        assert np.abs(ref["grad"][k]).max() > 0, k
        err = rel_l2(g[k], ref["grad"][k])
        print(f"[dft] {k}: rel L2 {err:.2e}", flush=True)
        assert err <= 1e-5, k


def test_dft_gradient_accumulates_and_switches(oracle_lib, gpu_ctx):
    """Two shots add up; nfreqs = 0 returns to backpropagation, which then
    equals the oracle's backpropagation gradient."""
    from seiscl_amd import Model
    csts, params, shot, true = _setup(dict(ND=2, FDORDER=4))
    sp, s, rp = shot
    sp2 = sp.copy()
    sp2[0] += 70.0
    dins = [oracle_lib.forward(true, p, s, rp, **csts) for p in (sp, sp2)]
    refs = [oracle_lib.gradient_dft(params, p, s, rp, FREQS, din=d, **csts)["grad"]
            for p, d in zip((sp, sp2), dins)]
    m = Model(gpu_ctx, params, **csts)
    m.grad_dft(FREQS)
    m.grad_zero()
    for p, d in zip((sp, sp2), dins):
        m.gradient(p, s, rp, din=d)
    g = m.grad_read()
    for k in refs[0]:
        assert rel_l2(g[k], refs[0][k] + refs[1][k]) <= 1e-5, k
    m.grad_dft(None)
    m.grad_zero()
    m.gradient(sp, s, rp, din=dins[0])
    g = m.grad_read()
    m.close()
    ref = oracle_lib.gradient(params, sp, s, rp, din=dins[0], **csts)["grad"]
    for k in ref:
        assert rel_l2(g[k], ref[k]) <= 1e-6, k


def test_dft_refusals(gpu_ctx):
    from seiscl_amd import Model
    from seiscl_amd._lib import SeisError
    csts, params, shot, _ = _setup(dict(ND=2, FDORDER=4))
    m = Model(gpu_ctx, params, **csts)
    with pytest.raises(SeisError, match="Nyquist"):
        m.grad_dft([600.0])
    with pytest.raises(SeisError, match="positive"):
        m.grad_dft([-1.0])
    with pytest.raises(SeisError, match="bin 0"):
        m.grad_dft([1.0])
    m.close()
    m = Model(gpu_ctx, params, **dict(csts, FP16=2))
    with pytest.raises(SeisError, match="FP32"):
        m.grad_dft(FREQS)
    m.close()
