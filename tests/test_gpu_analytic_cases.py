"""The HIP path on the cases that pin the oracle's explosive source, pressure
output, 2D acoustic update and free surface (tests/test_oracle_source_surface_pins.py):
the GPU equals the oracle on the same inputs and meets the closed forms
itself (Lamb's problem also at the reference's own 300^3 geometry); and the 3D viscoelastic benchmark geometry (explosion 10 cells deep,
receivers one cell below the free surface) on a reduced grid, past the
first arrival, in FP32, FP16 storage and FP64."""
import numpy as np
import pytest

from analytic import explosion_2d, explosion_3d, garvin_shot, lamb_shot
from cases import rel_l2
from seiscl_amd.workloads import config2, ricker

pytestmark = pytest.mark.gpu

VP, VS, RHO = 3500.0, 2000.0, 2000.0


def _homog(shape, acoustic=False):
    p = dict(vp=np.full(shape, VP, np.float32), rho=np.full(shape, RHO, np.float32))
    if not acoustic:
        p["vs"] = np.full(shape, VS, np.float32)
    return p


def _gpu(gpu_ctx, params, shot, **csts):
    from seiscl_amd import Model
    m = Model(gpu_ctx, params, **csts)
    out = m.forward(*shot)
    m.close()
    return out


def test_garvin_gpu(oracle_lib, gpu_ctx):
    N, dh, dt, NT, f0, FD, nab = 300, 2.0, 0.25e-3, 800, 20.0, 4, 112
    sx, sz = (nab + 2 * FD) * dh, 10 * dh
    gx = np.arange(sx + 5 * dh, (N - nab - 2 * FD) * dh, dh)
    shp = (N, N)
    sp = np.array([[sx], [0], [sz], [0], [100]])
    rp = np.stack([gx, 0 * gx, 0 * gx + sz, 0 * gx, np.arange(gx.size) + 1.0, 0 * gx, 0 * gx,
                   0 * gx])
    shot = (sp, ricker(f0, NT, dt)[:, None].astype(np.float32), rp)
    csts = dict(N=np.array(shp), ND=2, dh=dh, dt=dt, NT=NT, FDORDER=FD, abs_type=2, nab=nab,
                abpc=3.0, freesurf=1, seisout=1, f0=f0)
    print("[analytic] garvin 2D", flush=True)
    ref = oracle_lib.forward(_homog(shp), *shot, **csts)
    out = _gpu(gpu_ctx, _homog(shp), shot, **csts)
    for r, o in zip(ref, out):
        assert rel_l2(o, r) <= 1e-6
    fd = out[0].astype(np.float64)
    ana = garvin_shot(gx - sx + dh / 2, sz, sz, VP, VS, f0, dt, NT, comp="x")
    assert rel_l2(fd / np.linalg.norm(fd), ana / np.linalg.norm(ana)) < 6e-2


@pytest.mark.parametrize("comp", ["r", "t"])
def test_lamb_3d_gpu(oracle_lib, gpu_ctx, comp):
    from test_oracle_source_surface_pins import lamb_case
    p, shot, csts, k, ana = lamb_case(comp=comp)
    print(f"[analytic] lamb 3D {comp}", flush=True)
    ref = oracle_lib.forward(p, *shot, **csts)
    out = _gpu(gpu_ctx, p, shot, **csts)
    for r, o in zip(ref, out):
        assert rel_l2(o, r) <= 1e-6
    fd = out[k].astype(np.float64)
    assert rel_l2(fd / np.linalg.norm(fd), ana / np.linalg.norm(ana)) < 2e-2


@pytest.mark.parametrize("comp", ["r", "t"])
def test_lamb_3d_reference_geometry_gpu(gpu_ctx, comp):
    """SeisCL/tests/test_analytics.py:234-324 as written: 300^3, dh 2 m,
    dt 0.25 ms, NT 800, 20 Hz, FD4, CPML 112, force and receivers on the
    surface (inline: x; crossline: y), offsets 5 dh to the far absorbing
    layer, bar 6e-2 on unit-norm gathers."""
    from test_oracle_source_surface_pins import lamb_case
    p, shot, csts, k, ana = lamb_case(N=300, nab=112, NT=800, comp=comp)
    print(f"[analytic] lamb 3D reference geometry {comp}", flush=True)
    fd = _gpu(gpu_ctx, p, shot, **csts)[k].astype(np.float64)
    err = rel_l2(fd / np.linalg.norm(fd), ana / np.linalg.norm(ana))
    print(f"[analytic] lamb 3D 300^3 {comp}: {err:.3e}", flush=True)
    assert err < 6e-2, err


def test_explosion_3d_gpu(oracle_lib, gpu_ctx):
    N, dh, dt, NT, f0 = 90, 5.0, 0.5e-3, 400, 20.0
    c = N // 2
    shp = (N, N, N)
    sp = np.array([[c * dh], [c * dh], [c * dh], [0], [100]])
    ks = np.arange(c + 8, N - 22, 4)
    rp = np.stack([0 * ks + c * dh, 0 * ks + c * dh, ks * dh, 0 * ks, np.arange(ks.size) + 1.0,
                   0 * ks, 0 * ks, 0 * ks])
    shot = (sp, ricker(f0, NT, dt)[:, None].astype(np.float32), rp)
    csts = dict(N=np.array(shp), ND=3, dh=dh, dt=dt, NT=NT, FDORDER=4, abs_type=1, nab=20,
                f0=f0, VPPML=VP, FPML=f0)
    print("[analytic] explosion 3D", flush=True)
    ap, av = explosion_3d(VP, VS, RHO, dh, f0, dt, NT, (ks - c) * dh, (ks - c + 0.5) * dh)
    for so, ana, k in ((2, ap, 0), (1, av, 2)):
        ref = oracle_lib.forward(_homog(shp), *shot, seisout=so, nthreads=16, **csts)
        out = _gpu(gpu_ctx, _homog(shp), shot, seisout=so, **csts)
        for r, o in zip(ref, out):
            assert rel_l2(o, r) <= 1e-6
        assert rel_l2(out[k], ana) < 1e-2


@pytest.mark.parametrize("ND", [2, 22])
def test_explosion_2d_gpu(oracle_lib, gpu_ctx, ND):
    N, dh, dt, NT, f0 = 240, 5.0, 0.5e-3, 700, 20.0
    c = N // 2
    shp = (N, N)
    sp = np.array([[c * dh], [0], [c * dh], [0], [100 if ND == 2 else 3]])
    ks = np.arange(c + 8, N - 22, 4)
    rp = np.stack([0 * ks + c * dh, 0 * ks, ks * dh, 0 * ks, np.arange(ks.size) + 1.0,
                   0 * ks, 0 * ks, 0 * ks])
    shot = (sp, ricker(f0, NT, dt)[:, None].astype(np.float32), rp)
    csts = dict(N=np.array(shp), ND=ND, dh=dh, dt=dt, NT=NT, FDORDER=4, abs_type=1, nab=20,
                f0=f0, VPPML=VP, FPML=f0)
    m = _homog(shp, acoustic=ND == 22)
    print(f"[analytic] explosion 2D ND={ND}", flush=True)
    ap, av = explosion_2d(VP, VS, RHO, dh, f0, dt, NT, (ks - c) * dh, (ks - c + 0.5) * dh,
                          acoustic=ND == 22)
    for so, ana, k in ((2, ap, 0), (1, av, 1)):
        ref = oracle_lib.forward(m, *shot, seisout=so, **csts)
        out = _gpu(gpu_ctx, m, shot, seisout=so, **csts)
        for r, o in zip(ref, out):
            assert rel_l2(o, r) <= 1e-6
        assert rel_l2(out[k], ana) < 1e-2


def test_benchmark_geometry_precision(oracle_lib, gpu_ctx):
    """The 3D viscoelastic benchmark shot on a 64^3 grid: layered model,
    L=2 (Q=50 at 5-25 Hz), FD4, CPML 16, free surface, 12 Hz explosion 10
    cells deep, 16 x 16 receivers one cell deep, 400 steps (the first
    arrival reaches the receivers near step 200).  FP32 equals the oracle;
    against the FP64 run FP32 is within 1e-5 and FP16 storage within 1e-2."""
    csts, params, shots = config2(n=64, NT=400, nrec_side=16)
    print("[analytic] benchmark geometry 64^3", flush=True)
    ref = oracle_lib.forward(params, *shots[0], nthreads=16, **csts)
    o32 = _gpu(gpu_ctx, params, shots[0], **csts)
    o16 = _gpu(gpu_ctx, params, shots[0], **dict(csts, FP16=2))
    o64 = _gpu(gpu_ctx, params, shots[0], **dict(csts, FP16=64))
    for r, a, b, c in zip(ref, o32, o16, o64):
        assert np.abs(c).max() > 0
        assert rel_l2(a, r) <= 1e-6
        assert rel_l2(a, c) <= 1e-5
        assert rel_l2(b, c) <= 1e-2
