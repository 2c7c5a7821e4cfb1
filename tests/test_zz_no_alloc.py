"""seis_forward_shot_static: the forward shot that allocates no device memory
(its shot tables and seismograms live in the library's fixed per-GPU area)
gives the same seismograms, bit for bit, as seis_forward_shot, on two calls
in a row (the second with a smaller model and fewer receivers), and refuses
a shot that does not fit the area.  Called through the C-ABI (ctypes)."""
import ctypes

import numpy as np
import pytest

from cases import small_case
from seiscl_amd import Context, Model
from seiscl_amd.engine import fptr, shot_arrays

STATIC_BYTES = 64 << 20


def _static(m, shot):
    sp, s, rp = shot_arrays(*shot)
    out = np.full((len(m.out_names), rp.shape[0], m.csts.NT), np.nan, dtype=np.float32)
    rc = m.lib.seis_forward_shot_static(m.handle, sp.shape[0], fptr(sp), fptr(s), rp.shape[0],
                                        fptr(rp), fptr(out))
    return rc, out


CASES = [
    # (first call, second call with a smaller input, precision FP16=)
    (dict(ND=3, FDORDER=4, L=2, freesurf=1, NT=80),
     dict(ND=3, FDORDER=4, L=1, freesurf=1, NT=60, N=(32, 28, 36)), 0),
    (dict(ND=2, FDORDER=8, L=0, freesurf=0, NT=150),
     dict(ND=2, FDORDER=4, L=1, freesurf=1, NT=100, N=(48, 56)), 0),
    (dict(ND=3, FDORDER=4, L=2, freesurf=1, NT=80),
     dict(ND=3, FDORDER=4, L=0, freesurf=1, NT=60, N=(32, 28, 36)), 2),
    (dict(ND=2, FDORDER=4, L=1, freesurf=1, NT=100),
     dict(ND=3, FDORDER=4, L=0, freesurf=0, NT=60, N=(32, 28, 36)), 64),
]


@pytest.mark.gpu
@pytest.mark.parametrize("first,second,fp16", CASES)
def test_static_equals_forward_shot(first, second, fp16):
    ctx = Context(0)
    try:
        for k, case in enumerate((first, second)):
            csts, params, shot = small_case(**case)
            csts["FP16"] = fp16
            if k == 1:
                # fewer receivers on the smaller call
                rp = shot[2][:, ::2]
                shot = (shot[0], shot[1], rp)
            print(f"static shot call {k}: {case} FP16={fp16}", flush=True)
            m = Model(ctx, params, **csts)
            try:
                ref = m.forward(*shot, raw=True)
                rc, out = _static(m, shot)
                assert rc == 0, m.lib.seis_last_error().decode()
                np.testing.assert_array_equal(out, ref)
                assert np.abs(ref).max() > 0
                # the ordinary path still works after the static one
                again = m.forward(*shot, raw=True)
                np.testing.assert_array_equal(again, ref)
            finally:
                m.close()
    finally:
        ctx.close()


@pytest.mark.gpu
def test_static_refuses_oversize():
    ctx = Context(0)
    try:
        csts, params, shot = small_case(ND=2, NT=40)
        m = Model(ctx, params, **csts)
        try:
            sp, s, rp = shot
            # receivers x NT x outputs x 4 B x 2 (recording + transpose) > area
            nrec = STATIC_BYTES // (2 * 4 * 40 * len(m.out_names)) + 64
            rp_big = np.repeat(rp[:, :1], nrec, axis=1)
            rc, _ = _static(m, (sp, s, rp_big))
            assert rc != 0
            assert "static area" in m.lib.seis_last_error().decode()
        finally:
            m.close()
    finally:
        ctx.close()
