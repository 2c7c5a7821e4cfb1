"""Misfit options of the reference's residual computation (var_res_raw and
rtm_res, src/residuals.c:24-450; butterworth, src/butterworth.c:24-86):
the oracle against closed forms, numpy's FFT and finite differences, and the
library's host implementation (seis_misfit, no device needed) against the
oracle.

Pinned: the band-pass (exact-bin sinusoids are scaled by the filter's
H(f); arbitrary traces agree with numpy's rfft restatement of the same
periodic-extension filter), restype 1 (residuals = data), scaleshot (a
single shot-wide factor 1/|d|), scalermsnorm with the derivative fix
(finite differences of 1/2 sum |o/|o| - d/|d||^2), scalerms by a closed
form reduced by hand from the reference's steps (the reference states no
objective for it, and its source does not vanish at a perfect fit).
"""
import numpy as np
import pytest

from seiscl_amd import engine

BASE = dict(N=np.array([40, 50]), ND=2, dt=1e-3)


def _H(f, fmin, fmax, order=6):
    h = np.ones_like(f)
    if fmin > 0:
        h = h * (1 - np.sqrt(1 / (1 + (f / fmin) ** (2 * order))))
    if fmax > 0:
        h = h * np.sqrt(1 / (1 + (f / fmax) ** (2 * order)))
    return h


def _traces(NT=300, nr=5, nout=2, seed=0, noise=0.3):
    rng = np.random.default_rng(seed)
    out = [rng.standard_normal((NT, nr)).astype(np.float32) for _ in range(nout)]
    din = [(o + noise * rng.standard_normal(o.shape)).astype(np.float32) for o in out]
    return out, din


def test_butterworth_exact_bins(oracle_lib):
    NT, dt = 400, 1e-3                      # 400 = 2^4 5^2: no padding
    t = np.arange(NT) * dt
    for f in (5.0, 20.0, 35.0, 80.0, 150.0):
        x = np.cos(2 * np.pi * f * t + 0.3)[:, None].astype(np.float32)
        y = oracle_lib.butterworth(x, 20.0, 80.0, dt)
        np.testing.assert_allclose(y[:, 0], _H(f, 20.0, 80.0) * x[:, 0], atol=2e-6)


@pytest.mark.parametrize("tmax", [301, 250, 97])
def test_butterworth_matches_numpy(oracle_lib, tmax):
    NT, dt = 320, 1e-3
    rng = np.random.default_rng(tmax)
    x = rng.standard_normal((NT, 3)).astype(np.float32)
    y = oracle_lib.butterworth(x, 8.0, 60.0, dt, tmax=tmax)
    nfft = next(n for n in range(tmax, 4 * tmax) if set(_factors(n)) <= {2, 3, 5})
    ref = x.astype(np.float64).copy()
    for k in range(3):
        buf = np.concatenate([x[:tmax, k], x[:nfft - tmax, k]]).astype(np.float64)
        Y = np.fft.rfft(buf) * _H(np.arange(nfft // 2 + 1) / (dt * nfft), 8.0, 60.0)
        ref[:tmax, k] = np.fft.irfft(Y, nfft)[:tmax]
    np.testing.assert_allclose(y, ref, atol=2e-6)
    np.testing.assert_array_equal(y[tmax:], x[tmax:])


def _factors(n):
    f, p = [], 2
    while n > 1:
        while n % p == 0:
            f.append(p)
            n //= p
        p += 1
    return f


def test_plain_and_rtm(oracle_lib):
    out, din = _traces()
    res, (rms, rn) = oracle_lib.residuals(out, din, NT=300, seisout=1, **BASE)
    for r, o, d in zip(res, out, din):
        np.testing.assert_array_equal(r, o - d)
    assert rms == pytest.approx(sum(float(np.sum((o.astype(np.float64) - d) ** 2))
                                    for o, d in zip(out, din)), rel=1e-6)
    res, _ = oracle_lib.residuals(out, din, NT=300, seisout=1, restype=1, **BASE)
    for r, d in zip(res, din):
        np.testing.assert_array_equal(r, d)


def test_scaleshot_closed_form(oracle_lib):
    out, din = _traces()
    res, _ = oracle_lib.residuals(out, din, NT=300, seisout=1, scaleshot=1, **BASE)
    ws = 1 / np.sqrt(sum(float(np.sum(d.astype(np.float64) ** 2)) for d in din))
    for r, o, d in zip(res, out, din):
        np.testing.assert_allclose(r, ws * ws * (o.astype(np.float64) - d), rtol=2e-5,
                                   atol=1e-7 * ws * ws)


def test_scalermsnorm_is_the_derivative(oracle_lib):
    """With the fix the adjoint source is the gradient of
    J = 1/2 sum_g |o_g/|o_g| - d_g/|d_g||^2; the reference's form is not."""
    out, din = _traces(NT=200, nr=3, nout=1, seed=1, noise=0.5)
    o, d = out[0].astype(np.float64), din[0].astype(np.float64)
    kw = dict(NT=200, seisout=2, scalermsnorm=1, **BASE)
    fixed = oracle_lib.residuals(out, din, **kw)[0][0].astype(np.float64)
    ref = oracle_lib.residuals(out, din, gradfix=0, **kw)[0][0].astype(np.float64)

    def J(x):
        xn = x / np.linalg.norm(x, axis=0)
        dn = d / np.linalg.norm(d, axis=0)
        return 0.5 * np.sum((xn - dn) ** 2)

    h = 1e-6
    for t in range(3, 200, 29):
        for g in range(3):
            e = np.zeros_like(o)
            e[t, g] = h
            fd = (J(o + e) - J(o - e)) / (2 * h)
            assert fixed[t, g] == pytest.approx(fd, rel=1e-4, abs=1e-7)
    assert np.linalg.norm(ref - fixed) > 0.1 * np.linalg.norm(fixed)


def test_scalerms_closed_form(oracle_lib):
    """scalerms (src/residuals.c:123-186, 209-237, 327-341) reduced by hand
    to a closed form in double.  With per-receiver norms |o_g|, |d_g| over
    all outputs and samples and the shot norm |D|: the outputs are scaled by
    c_g = |d_g| / (|o_g| |D|) and the data by 1/|D|; the factor that
    multiplies the normalised misfit is |D|^2 (the data scale is overwritten
    by 1/|D| before it is read), so the source is
    c_g (c_g (|d_g|^2 - |d_g| <o_g, d_g> / |o_g|) o_g - d_g / |D|),
    the scaled data have unit energy (rmsnorm = 1) and every scaled output
    trace has its scaled data trace's energy."""
    out, din = _traces(NT=300, nr=4, nout=2, seed=3, noise=0.4)
    res, (rms, rmsnorm) = oracle_lib.residuals(out, din, NT=300, seisout=1, scalerms=1, **BASE)
    o = np.stack([x.astype(np.float64) for x in out])          # (nout, NT, nr)
    d = np.stack([x.astype(np.float64) for x in din])
    no = np.sqrt(np.sum(o ** 2, axis=(0, 1)))
    nd = np.sqrt(np.sum(d ** 2, axis=(0, 1)))
    nD = np.sqrt(np.sum(d ** 2))
    c = nd / (no * nD)
    inner = np.sum(o * d, axis=(0, 1))
    want = c * (c * (nd ** 2 - nd * inner / no) * o - d / nD)
    got = np.stack(res).astype(np.float64)
    assert np.linalg.norm(got - want) <= 1e-4 * np.linalg.norm(want)   # float sums: 7e-6
    assert rmsnorm == pytest.approx(1.0, rel=1e-5)
    assert rms == pytest.approx(float(np.sum((d / nD - c * o) ** 2)), rel=1e-4)


def test_scalerms_reference_form(oracle_lib):
    """Unpinned beyond library == oracle: at a perfect fit (o = d) the
    reference's scalerms source is -ws^2 d, not zero."""
    out, _ = _traces(nout=1)
    res, _ = oracle_lib.residuals(out, out, NT=300, seisout=2, scalerms=1, **BASE)
    ws = 1 / np.sqrt(float(np.sum(out[0].astype(np.float64) ** 2)))
    np.testing.assert_allclose(res[0], -ws * ws * out[0], rtol=1e-4, atol=1e-9)


OPTIONS = [dict(), dict(scalerms=1), dict(scaleshot=1), dict(scalermsnorm=1),
           dict(scalermsnorm=1, gradfix=0), dict(fmin=5.0, fmax=60.0), dict(fmax=40.0),
           dict(restype=1), dict(scalerms=1, fmin=5.0, fmax=60.0, tmin=0.02, tmax=0.25),
           dict(scaleshot=1, fmin=10.0, tmax=0.199)]


@pytest.mark.parametrize("opt", OPTIONS, ids=lambda o: "-".join(f"{k}{v}" for k, v in o.items())
                         or "plain")
def test_library_misfit_equals_oracle(oracle_lib, opt):
    out, din = _traces()
    opt = dict(opt)
    gradfix = opt.pop("gradfix", oracle_lib.FIX_ALL)
    kw = dict(NT=300, seisout=1, **BASE, **opt)
    a, ra = oracle_lib.residuals(out, din, gradfix=gradfix, **kw)
    b, rb = engine.misfit(out, din, gradfix=gradfix, **kw)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert ra == rb
