"""Seeded synthetic workloads of the benchmark configurations.

Each builder returns ``(csts, params, shots)`` where ``csts`` holds
SeisCL.py-style constants, ``params`` the material arrays ([NZ, NX] or
[NZ, NY, NX]) and ``shots`` a list of ``(src_pos, src, rec_pos)`` triplets in
SeisCL.py shapes ((5, ns), (NT, ns), (8, nr)).

Smooth random fields are white noise passed through a Gaussian filter of
sigma = correlation length / dh cells, normalised to max |.| = 1.
"""
from __future__ import annotations

import numpy as np
from scipy.ndimage import gaussian_filter


def ricker(f0: float, NT: int, dt: float, tmin: float | None = None) -> np.ndarray:
    """Ricker wavelet of SeisCL.ricker_wavelet (SeisCL/SeisCL.py:810-834)."""
    if tmin is None:
        tmin = -1.5 / f0
    t = np.linspace(tmin, (NT - 1) * dt + tmin, num=int(NT))
    return ((1.0 - 2.0 * (np.pi ** 2) * (f0 ** 2) * (t ** 2))
            * np.exp(-(np.pi ** 2) * (f0 ** 2) * (t ** 2)))


def smooth(shape, corr_len: float, dh: float, rng) -> np.ndarray:
    g = gaussian_filter(rng.standard_normal(shape), sigma=corr_len / dh, mode="nearest")
    return g / np.abs(g).max()


def q_to_tau(Q: float, fmin: float, fmax: float, FL, Qmin: float = 5.0,
             Qmax: float = 500.0) -> float:
    """tau giving an average 1/Q over [fmin, fmax] for relaxation frequencies
    FL: QTAU(Qmin, Qmax, fmin, fmax, FL).Q2tau(Q) (SeisCL/Q_tau.py:25-131)."""
    from .qtau import QTAU
    return float(QTAU(Qmin, Qmax, fmin, fmax, FL).Q2tau(Q))


def layered(shape, nlayers: int, vmin: float, vmax: float, rng, axis: int = 0) -> np.ndarray:
    """1D layered velocity increasing with depth (axis 0 = z), seeded interfaces."""
    nz = shape[axis]
    cuts = np.sort(rng.choice(np.arange(1, nz), size=nlayers - 1, replace=False))
    v1 = np.zeros(nz)
    vals = np.linspace(vmin, vmax, nlayers)
    edges = np.concatenate([[0], cuts, [nz]])
    for k in range(nlayers):
        v1[edges[k]:edges[k + 1]] = vals[k]
    sh = [1] * len(shape)
    sh[axis] = nz
    return np.broadcast_to(v1.reshape(sh), shape).copy()


def config1(NT: int = 2000, seed: int = 1):
    """Plain 2D isotropic elastic, one shot: 200 x 400, dh 10 m, dt 1 ms,
    FD4, CPML 20, no free surface, vz Ricker 10 Hz at (50 m, 2000 m),
    400 receivers at z = 50 m."""
    rng = np.random.default_rng(seed)
    N = (200, 400)
    dh, dt = 10.0, 1e-3
    vp = 3500 * (1 + 0.05 * smooth(N, 100, dh, rng))
    vs = 2000 * (1 + 0.05 * smooth(N, 100, dh, rng))
    rho = 2000 * (1 + 0.05 * smooth(N, 100, dh, rng))
    csts = dict(N=np.array(N), ND=2, dh=dh, dt=dt, NT=NT, FDORDER=4, MAXRELERROR=1,
                abs_type=1, nab=20, freesurf=0, L=0, f0=10.0, seisout=1)
    src_pos = np.array([[2000.0], [0.0], [50.0], [0.0], [2.0]])
    src = ricker(10.0, NT, dt)[:, None]
    gx = np.arange(400) * dh
    rec_pos = np.stack([gx, 0 * gx, 0 * gx + 50.0, 0 * gx, np.arange(400) + 1.0,
                        0 * gx, 0 * gx, 0 * gx])
    params = dict(vp=vp.astype(np.float32), vs=vs.astype(np.float32),
                  rho=rho.astype(np.float32))
    return csts, params, [(src_pos, src, rec_pos)]


def config2(n: int = 256, NT: int = 1500, seed: int = 2, nrec_side: int = 64,
            elastic: bool = False):
    """Plain 3D viscoelastic, one shot: n^3, dh 10 m, dt 0.8 ms, FD4, CPML 16,
    free surface, L=2 (FL 5 and 25 Hz, Q=50), layered vp 1800-4500,
    vs = vp/1.8, rho 1800-2600, 3% smooth perturbations, 12 Hz Ricker
    explosion, 64 x 64 surface receivers.  elastic=True: the same shot and
    model without attenuation (L=0)."""
    rng = np.random.default_rng(seed)
    N = (n, n, n)
    dh, dt = 10.0, 8e-4
    vp = layered(N, 8, 1800.0, 4500.0, rng)
    pert = smooth(N, 100, dh, rng)
    vp = vp * (1 + 0.03 * pert)
    vs = vp / 1.8
    rho = 1800 + 800 * (vp - 1800) / 2700
    FL = [5.0, 25.0]
    tau = q_to_tau(50.0, 5.0, 25.0, FL)
    taup = np.full(N, tau, np.float32)
    taus = np.full(N, tau, np.float32)
    csts = dict(N=np.array(N), ND=3, dh=dh, dt=dt, NT=NT, FDORDER=4, MAXRELERROR=1,
                abs_type=1, nab=16, freesurf=1, L=2, FL=np.array(FL), f0=12.0, seisout=1)
    c = n // 2 * dh
    src_pos = np.array([[c], [c], [10 * dh], [0.0], [100.0]])
    src = ricker(12.0, NT, dt)[:, None]
    k = min(nrec_side, n // 3)
    off = (n - 3 * (k - 1)) // 2
    gi = (off + 3 * np.arange(k)) * dh
    gx, gy = np.meshgrid(gi, gi, indexing="ij")
    gx, gy = gx.ravel(), gy.ravel()
    nr = gx.size
    rec_pos = np.stack([gx, gy, 0 * gx + dh, 0 * gx, np.arange(nr) + 1.0, 0 * gx,
                        0 * gx, 0 * gx])
    params = dict(vp=vp.astype(np.float32), vs=vs.astype(np.float32),
                  rho=rho.astype(np.float32), taup=taup, taus=taus)
    if elastic:
        csts["L"] = 0
        del params["taup"], params["taus"]
    return csts, params, [(src_pos, src, rec_pos)]


def config5(n: int = 512, NT: int = 2500, nshots: int = 8, seed: int = 5,
            nrec_side: int = 64):
    """3D viscoelastic FWI gradient batch: n^3 (512), dh 10 m, dt 0.8 ms,
    NT 2500, FD4, CPML 20, free surface, L=2 (FL 5 and 25 Hz, Q=50), model
    as config2 (layered vp 1800-4500, vs = vp/1.8, rho 1800-2600, 3% smooth
    perturbations at 100 m, drawn in float32 at this size), 8 explosive
    12 Hz Ricker shots on a 4 x 2 grid at z = 10 cells, 64 x 64 surface
    receivers.  FP16 wavefield storage is the caller's choice (csts FP16)."""
    rng = np.random.default_rng(seed)
    N = (n, n, n)
    dh, dt = 10.0, 8e-4
    vp = layered(N, 8, 1800.0, 4500.0, rng).astype(np.float32)
    noise = rng.standard_normal(N, dtype=np.float32)
    pert = gaussian_filter(noise, sigma=100.0 / dh, mode="nearest")
    del noise
    pert /= np.abs(pert).max()
    vp *= 1 + np.float32(0.03) * pert
    del pert
    FL = [5.0, 25.0]
    tau = np.float32(q_to_tau(50.0, 5.0, 25.0, FL))
    params = dict(vp=vp, vs=vp / np.float32(1.8),
                  rho=np.float32(1800) + np.float32(800 / 2700) * (vp - np.float32(1800)),
                  taup=np.full(N, tau, np.float32), taus=np.full(N, tau, np.float32))
    csts = dict(N=np.array(N), ND=3, dh=dh, dt=dt, NT=NT, FDORDER=4, MAXRELERROR=1,
                abs_type=1, nab=20, freesurf=1, L=2, FL=np.array(FL), f0=12.0, seisout=1)
    k = min(nrec_side, n // 3)
    off = (n - 3 * (k - 1)) // 2
    gi = (off + 3 * np.arange(k)) * dh
    gx, gy = np.meshgrid(gi, gi, indexing="ij")
    gx, gy = gx.ravel(), gy.ravel()
    src = ricker(12.0, NT, dt)[:, None].astype(np.float32)
    shots = []
    xs = np.linspace(0.2, 0.8, 4) * n * dh
    ys = np.linspace(0.35, 0.65, 2) * n * dh
    for j, (x, y) in enumerate((x, y) for y in ys for x in xs):
        if j >= nshots:
            break
        sp = np.array([[x], [y], [10 * dh], [float(j)], [100.0]])
        rp = np.stack([gx, gy, 0 * gx + dh, 0 * gx + j, j * gx.size + np.arange(gx.size) + 1.0,
                       0 * gx, 0 * gx, 0 * gx])
        shots.append((sp, src, rp))
    return csts, params, shots


def layered_model(shape, rng, nlayers: int = 6, vmin: float = 1800.0, vmax: float = 4000.0,
                  elastic: bool = True):
    """Seeded layered model: vp from vmin to vmax with depth, vs = vp/1.8,
    rho = 1800 + 800 (vp - 1800) / 2700."""
    vp = layered(shape, nlayers, vmin, vmax, rng)
    rho = 1800 + 800 * (vp - 1800) / 2700
    p = dict(vp=vp.astype(np.float32), rho=rho.astype(np.float32))
    if elastic:
        p["vs"] = (vp / 1.8).astype(np.float32)
    return p


def config3(N=(300, 800), NT: int = 3000, nshots: int = 16, seed: int = 3, f0: float = 8.0):
    """2D elastic FWI gradient: 300 x 800, dh 10 m, dt 1 ms, NT 3000, FD8,
    CPML 16.  True model: 6 seeded layers vp 1800-4000, vs = vp/1.8,
    rho = 1800 + 800 (vp - 1800) / 2700, plus 3 Gaussian anomalies (+-10%,
    radius 150 m); starting model: the true model smoothed with a 200 m
    Gaussian.  16 explosive shots evenly spaced along x and receivers every
    2 cells, both at z = nab + 5 cells; Ricker f0 (8 Hz: 5 cells per shortest
    S wavelength at 2.5 f0).  Returns (csts, true params, start params, shots)."""
    rng = np.random.default_rng(seed)
    NZ, NX = N
    dh, dt, nab = 10.0, 1e-3, 16
    vp = layered(N, 6, 1800.0, 4000.0, rng)
    zz, xx = np.meshgrid(np.arange(NZ) * dh, np.arange(NX) * dh, indexing="ij")
    for _ in range(3):
        cz = rng.uniform(0.3, 0.8) * NZ * dh
        cx = rng.uniform(0.2, 0.8) * NX * dh
        amp = rng.choice([-0.1, 0.1])
        vp = vp * (1 + amp * np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / 150.0 ** 2))
    true = dict(vp=vp.astype(np.float32), vs=(vp / 1.8).astype(np.float32),
                rho=(1800 + 800 * (vp - 1800) / 2700).astype(np.float32))
    start = {k: gaussian_filter(v.astype(np.float64), sigma=200.0 / dh,
                                mode="nearest").astype(np.float32) for k, v in true.items()}
    csts = dict(N=np.array(N), ND=2, dh=dh, dt=dt, NT=NT, FDORDER=8, MAXRELERROR=1,
                abs_type=1, nab=nab, freesurf=0, L=0, f0=f0, seisout=1)
    zs = (nab + 5) * dh
    src = ricker(f0, NT, dt)[:, None].astype(np.float32)
    gx = np.arange(nab, NX - nab, 2) * dh
    shots = []
    for k, x in enumerate(np.linspace(nab + 5, NX - nab - 5, nshots).round() * dh):
        sp = np.array([[x], [0.0], [zs], [float(k)], [100.0]])
        rp = np.stack([gx, 0 * gx, 0 * gx + zs, 0 * gx + k, k * gx.size + np.arange(gx.size) + 1.0,
                       0 * gx, 0 * gx, 0 * gx])
        shots.append((sp, src, rp))
    return csts, true, start, shots


def config4_request(rng, nab: int = 16, NZ=None, NX=None, NT=None):
    """One serving request: 2D acoustic (ND 22) or elastic (ND 2) modeling,
    NZ ~ U{100..600} even, NX ~ U{200..1500}, NT ~ U{1000..5000}, FD 4 or
    8, dh 10 m, dt 1 ms, Ricker 5-20 Hz, layered model, one source (acoustic
    type 3, elastic explosion) and receivers every 2 cells at z = nab+2."""
    NZ = NZ or 2 * int(rng.integers(50, 301))
    NX = NX or int(rng.integers(200, 1501))
    NT = NT or int(rng.integers(1000, 5001))
    fd = int(rng.choice([4, 8]))
    ND = int(rng.choice([22, 2]))
    f0 = float(rng.uniform(5.0, 20.0))
    dh, dt = 10.0, 1e-3
    params = layered_model((NZ, NX), rng, elastic=ND == 2)
    csts = dict(N=np.array([NZ, NX]), ND=ND, dh=dh, dt=dt, NT=NT, FDORDER=fd, MAXRELERROR=1,
                abs_type=1, nab=nab, freesurf=0, L=0, f0=f0, seisout=2 if ND == 22 else 1)
    sx = float(rng.integers(nab + 5, NX - nab - 5)) * dh
    src_pos = np.array([[sx], [0.0], [(nab + 2) * dh], [0.0], [3.0 if ND == 22 else 100.0]])
    src = ricker(f0, NT, dt)[:, None]
    gx = np.arange(nab, NX - nab, 2) * dh
    rec_pos = np.stack([gx, 0 * gx, 0 * gx + (nab + 2) * dh, 0 * gx, np.arange(gx.size) + 1.0,
                        0 * gx, 0 * gx, 0 * gx])
    return csts, params, (src_pos, src.astype(np.float32), rec_pos)


def config4(n: int = 256, seed: int = 4, **kw):
    """The serving workload: n seeded requests (config4_request)."""
    rng = np.random.default_rng(seed)
    return [config4_request(rng, **kw) for _ in range(n)]
