"""Precision modes of the GPU kernels (seis_csts.FP16), 3D, 2D P-SV and 2D
acoustic.

  FP16=1  FP32 storage in the scaled units used for FP16: the scaling is by
          powers of two, so results must equal the FP32 oracle (up to
          subnormals) -- this pins the scaling of coefficients, sources and
          outputs independently of FP16 rounding;
  FP16=64 FP64 storage and arithmetic with the same coefficients: the FP32
          path must agree with it to 1e-5 (relative L2 on seismograms);
  FP16=2  FP16 storage of wavefields and memory variables, FP32 arithmetic:
          within 1e-2 of the FP64 run.
"""
import numpy as np
import pytest

from cases import rel_l2, small_case

pytestmark = pytest.mark.gpu

CASES = ([dict(ND=3, FDORDER=o, abs_type=a, freesurf=f, L=l)
          for o in (4, 8) for a in (1, 2) for f in (0, 1) for l in (0, 2)]
         # several z tiles tall
         + [dict(ND=3, FDORDER=4, abs_type=1, freesurf=1, L=2, N=(150, 28, 30))]
         + [dict(ND=3, FDORDER=o, abs_type=1, freesurf=1, L=l, N=(128, 20, 22))
            for o in (4, 8) for l in (0, 2)]
         # 2D P-SV (src/update_{v,s}2D_half2.cl) and 2D acoustic
         + [dict(ND=2, FDORDER=o, abs_type=a, freesurf=f, L=l)
            for o in (4, 8) for a in (1, 2) for f in (0, 1) for l in (0, 2)]
         + [dict(ND=22, FDORDER=o, abs_type=a, L=l) for o, a, l in ((4, 1, 0), (8, 2, 1))])


def _ids(c):
    return "-".join(f"{k}{v}" for k, v in c.items())


def _run(gpu_ctx, csts, params, shot, fp16, fields=()):
    from seiscl_amd import Model
    m = Model(gpu_ctx, params, **dict(csts, FP16=fp16))
    out = m.forward(*shot)
    f = {k: m.read_field(k) for k in fields}
    m.close()
    return out, f


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_scaled_units_equal_fp32(oracle_lib, gpu_ctx, case):
    csts, params, shot = small_case(NT=100 if case["ND"] == 3 else 150, **case)
    print(f"[precision] scaled case={_ids(case)}", flush=True)
    ref, rf = oracle_lib.forward(params, *shot, fields=True, **csts)
    fields = ("vx", "p") if case["ND"] == 22 else ("vx", "szz")
    out, f = _run(gpu_ctx, csts, params, shot, 1, fields=fields)
    for r, o in zip(ref, out):
        assert rel_l2(o, r) <= 1e-6
    for k in f:
        assert rel_l2(f[k], rf[k]) <= 1e-6, k


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_fp32_and_fp16_against_fp64(gpu_ctx, case):
    csts, params, shot = small_case(NT=120 if case["ND"] == 3 else 150, **case)
    print(f"[precision] fp64/fp16 case={_ids(case)}", flush=True)
    o64, _ = _run(gpu_ctx, csts, params, shot, 64)
    o32, _ = _run(gpu_ctx, csts, params, shot, 0)
    o16, f16 = _run(gpu_ctx, csts, params, shot, 2, fields=("vz",))
    for a, b, c in zip(o32, o16, o64):
        assert np.abs(c).max() > 0
        assert rel_l2(a, c) <= 1e-5
        assert rel_l2(b, c) <= 1e-2
        assert rel_l2(b, c) > rel_l2(a, c)      # FP16 storage is really in use
    assert np.isfinite(f16["vz"]).all()


@pytest.mark.parametrize("case", [dict(ND=2, FDORDER=4, freesurf=1), dict(ND=2, FDORDER=8, L=2),
                                  dict(ND=22, FDORDER=4, abs_type=2), dict(ND=22, FDORDER=8, L=1)],
                         ids=_ids)
def test_fp64_2d(gpu_ctx, case):
    """2D runs in FP64 (used for finite-difference checks of gradients):
    the FP32 path agrees with them to 1e-5."""
    csts, params, shot = small_case(NT=150, **case)
    o64, _ = _run(gpu_ctx, csts, params, shot, 64)
    o32, _ = _run(gpu_ctx, csts, params, shot, 0)
    for a, c in zip(o32, o64):
        assert np.abs(c).max() > 0
        assert 0 < rel_l2(a, c) <= 1e-5
