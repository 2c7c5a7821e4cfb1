"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker used by tests/, the smoke entry point
and bench.py's cpu_baseline leg.  The product package never imports it.

Calling conventions match seiscl_amd.engine (SeisCL.py array shapes).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
MAXL = 8

DEFAULTS = dict(
    ND=2, dh=10.0, dt=0.0008, NT=875, L=0, f0=15.0, FL=(15.0,), FDORDER=8,
    MAXRELERROR=1, freesurf=0, abs_type=1, VPPML=3500.0, NPOWER=2.0, FPML=15.0,
    K_MAX_CPML=2.0, nab=16, abpc=6.0, seisout=2, param_type=0, tmax=0.0, tmin=0.0,
    restype=0, scalerms=0, scaleshot=0, scalermsnorm=0, fmin=0.0, fmax=0.0,
)


class OrCsts(ctypes.Structure):
    _fields_ = [
        ("ND", ctypes.c_int), ("N", ctypes.c_int * 3), ("NT", ctypes.c_int),
        ("tmax", ctypes.c_int), ("tmin", ctypes.c_int),
        ("dt", ctypes.c_float), ("dh", ctypes.c_float),
        ("FDORDER", ctypes.c_int), ("MAXRELERROR", ctypes.c_int),
        ("freesurf", ctypes.c_int), ("abs_type", ctypes.c_int), ("nab", ctypes.c_int),
        ("abpc", ctypes.c_float), ("VPPML", ctypes.c_float), ("NPOWER", ctypes.c_float),
        ("FPML", ctypes.c_float), ("K_MAX_CPML", ctypes.c_float),
        ("L", ctypes.c_int), ("f0", ctypes.c_float), ("FL", ctypes.c_float * MAXL),
        ("seisout", ctypes.c_int), ("param_type", ctypes.c_int), ("nthreads", ctypes.c_int),
        ("restype", ctypes.c_int), ("scalerms", ctypes.c_int), ("scaleshot", ctypes.c_int),
        ("scalermsnorm", ctypes.c_int), ("fmin", ctypes.c_float), ("fmax", ctypes.c_float),
    ]


_FP = ctypes.POINTER(ctypes.c_float)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        lib.or_forward.restype = ctypes.c_int
        lib.or_forward.argtypes = [ctypes.POINTER(OrCsts), _FP, _FP, _FP, _FP, _FP,
                                   ctypes.c_int, _FP, _FP, ctypes.c_int, _FP, _FP, _FP,
                                   ctypes.c_char_p, ctypes.c_int]
        lib.or_forward_from.restype = ctypes.c_int
        lib.or_forward_from.argtypes = [ctypes.POINTER(OrCsts), _FP, _FP, _FP, _FP, _FP,
                                        ctypes.c_int, _FP, _FP, ctypes.c_int, _FP, _FP, _FP,
                                        _FP, ctypes.c_char_p, ctypes.c_int]
        lib.or_last_loop_seconds.restype = ctypes.c_double
        lib.or_last_loop_seconds.argtypes = []
        lib.or_holberg.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        lib.or_kernel_hc.argtypes = [ctypes.c_int, ctypes.c_int, _FP]
        lib.or_literal9.restype = ctypes.c_float
        lib.or_literal9.argtypes = [ctypes.c_double]
        lib.or_cpml_coeff.argtypes = [ctypes.c_float] * 6 + [ctypes.c_int] + [_FP] * 6
        lib.or_taper.argtypes = [ctypes.c_int, ctypes.c_float, _FP]
        lib.or_eta.argtypes = [ctypes.c_int, _FP, ctypes.c_float, _FP]
        lib.or_transform.argtypes = [ctypes.POINTER(OrCsts)] + [_FP] * 19
        lib.or_check_stability.argtypes = [ctypes.POINTER(OrCsts), _FP, _FP, _FP, _FP]
        lib.or_nout.argtypes = [ctypes.POINTER(OrCsts)]
        lib.or_outname.restype = ctypes.c_char_p
        lib.or_outname.argtypes = [ctypes.POINTER(OrCsts), ctypes.c_int]
        lib.or_gradient.restype = ctypes.c_int
        lib.or_gradient.argtypes = ([ctypes.POINTER(OrCsts)] + [_FP] * 5 +
                                    [ctypes.c_int, _FP, _FP, ctypes.c_int, _FP, _FP, _FP,
                                     ctypes.c_int, _FP, _FP, _FP,
                                     ctypes.POINTER(ctypes.c_double), _FP, _FP, _FP,
                                     ctypes.c_int, _FP, ctypes.c_char_p, ctypes.c_int])
        _D = ctypes.POINTER(ctypes.c_double)
        lib.or_gradient_dft.restype = ctypes.c_int
        lib.or_gradient_dft.argtypes = ([ctypes.POINTER(OrCsts)] + [_FP] * 5 +
                                        [ctypes.c_int, _FP, _FP, ctypes.c_int, _FP, _FP, _FP,
                                         ctypes.c_int, ctypes.c_int, _FP, _FP, _D, _D, _FP,
                                         ctypes.c_char_p, ctypes.c_int])
        lib.or_dft_sampling.restype = ctypes.c_int
        lib.or_dft_sampling.argtypes = [ctypes.POINTER(OrCsts), ctypes.c_int, _FP,
                                        ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int), _FP, _FP]
        lib.or_dft_coef.restype = None
        lib.or_dft_coef.argtypes = [_D, ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 8
        lib.or_shell_check.restype = ctypes.c_int
        lib.or_shell_check.argtypes = [ctypes.POINTER(OrCsts), ctypes.POINTER(ctypes.c_long)]
        lib.or_nvars.argtypes = [ctypes.POINTER(OrCsts)]
        lib.or_varname.restype = ctypes.c_char_p
        lib.or_varname.argtypes = [ctypes.POINTER(OrCsts), ctypes.c_int]
        _lib = lib
    return _lib


def fp(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_FP)


def make_csts(nthreads=0, **kw) -> OrCsts:
    d = dict(DEFAULTS)
    d.update(kw)
    c = OrCsts()
    c.ND = int(d["ND"])
    for i, v in enumerate(np.atleast_1d(d["N"])):
        c.N[i] = int(v)
    c.NT = int(d["NT"])
    c.dt = float(d["dt"])
    c.dh = float(d["dh"])
    c.tmax = int(np.float32(d["tmax"]) / np.float32(c.dt)) if d["tmax"] else 0
    c.tmin = int(np.float32(d["tmin"]) / np.float32(c.dt)) if d["tmin"] else 0
    for k in ("FDORDER", "MAXRELERROR", "freesurf", "abs_type", "nab", "L", "seisout",
              "param_type", "restype", "scalerms", "scaleshot", "scalermsnorm"):
        setattr(c, k, int(d[k]))
    for k in ("abpc", "VPPML", "NPOWER", "FPML", "K_MAX_CPML", "f0", "fmin", "fmax"):
        setattr(c, k, float(d[k]))
    fl = np.atleast_1d(np.asarray(d["FL"], dtype=np.float32))
    for i in range(min(len(fl), MAXL)):
        c.FL[i] = float(fl[i])
    c.nthreads = int(nthreads)
    return c


def residuals(out, din, gradfix=None, **csts):
    """Oracle residuals of one shot (var_res_raw with the misfit options, or
    rtm_res): lists of (NT, nrec) arrays in, (res list, (rms, rmsnorm)) out."""
    lib = load()
    c = make_csts(**csts)
    o = np.ascontiguousarray(np.stack([np.asarray(x, np.float32).T for x in out]))
    d = np.ascontiguousarray(np.stack([np.asarray(x, np.float32).T for x in din]))
    r = np.zeros_like(o)
    rms = (ctypes.c_double * 2)()
    lib.or_residuals(ctypes.byref(c), o.shape[0], o.shape[1], fp(o), fp(d), fp(r), rms,
                     int(FIX_ALL if gradfix is None else gradfix))
    return [np.ascontiguousarray(x.T) for x in r], (rms[0], rms[1])


def butterworth(traces, fmin, fmax, dt, order=6, tmax=None):
    """Oracle band-pass of (NT, ntrace) traces (a filtered copy)."""
    lib = load()
    a = np.ascontiguousarray(np.asarray(traces, np.float32).T).copy()
    NT, ntr = a.shape[1], a.shape[0]
    lib.or_butterworth(fp(a), ctypes.c_float(fmin), ctypes.c_float(fmax), ctypes.c_float(dt),
                       NT, int(tmax or NT), ntr, int(order))
    return np.ascontiguousarray(a.T)


def to_c(a):
    return None if a is None else np.ascontiguousarray(np.asarray(a, np.float32).T).ravel()


def out_names(c: OrCsts):
    lib = load()
    return [lib.or_outname(ctypes.byref(c), i).decode() for i in range(lib.or_nout(ctypes.byref(c)))]


def var_names(c: OrCsts):
    lib = load()
    return [lib.or_varname(ctypes.byref(c), i).decode() for i in range(lib.or_nvars(ctypes.byref(c)))]


def _params(params, c: OrCsts):
    pt = c.param_type
    names = {0: ("vp", "vs"), 1: ("M", "mu"), 2: ("Ip", "Is"), 3: ("vpR", "vsR")}[pt]
    p0 = to_c(params[names[0]])
    p1 = to_c(params[names[1]]) if c.ND != 22 else None
    rho = to_c(params["rho"])
    taup = taus = None
    if c.L > 0:
        taup = to_c(params["vpI" if pt == 3 else "taup"])
        if c.ND != 22:
            taus = to_c(params["vsI" if pt == 3 else "taus"])
    return p0, p1, rho, taup, taus


def forward(params: dict, src_pos, src, rec_pos, nthreads: int = 0, fields: bool = False,
            fields_in: dict | None = None, **csts):
    """Oracle forward modeling of one shot.

    Returns a list of (NT, nrec) seismograms in out_names order, plus a dict of
    final padded wavefields ([NZp, (NYp,) NXp]) when fields=True.  fields_in
    (same layout, every wavefield named) starts the run from those fields
    instead of zeros."""
    lib = load()
    c = make_csts(nthreads=nthreads, **csts)
    p0, p1, rho, taup, taus = _params(params, c)
    sp = np.ascontiguousarray(np.asarray(src_pos, np.float32).T)
    s = np.ascontiguousarray(np.asarray(src, np.float32).T)
    rp = np.ascontiguousarray(np.asarray(rec_pos, np.float32).T)
    nout = lib.or_nout(ctypes.byref(c))
    out = np.zeros((nout, rp.shape[0], c.NT), np.float32)
    shape = [int(v) for v in np.atleast_1d(csts["N"])]
    H = c.FDORDER // 2
    pshape = [n + 2 * H for n in shape]
    nvars = lib.or_nvars(ctypes.byref(c))
    fo = np.zeros(nvars * int(np.prod(pshape)), np.float32) if fields else None
    fi = None
    if fields_in is not None:
        fi = np.concatenate([np.ascontiguousarray(np.asarray(fields_in[name], np.float32).T).ravel()
                             for name in var_names(c)])
    err = ctypes.create_string_buffer(256)
    st = lib.or_forward_from(ctypes.byref(c), fp(p0), fp(p1), fp(rho), fp(taup), fp(taus),
                             sp.shape[0], fp(sp), fp(s), rp.shape[0], fp(rp), fp(fi), fp(out),
                             fp(fo), err, 256)
    if st:
        raise RuntimeError("oracle: " + err.value.decode())
    seis = [np.ascontiguousarray(o.T) for o in out]
    if not fields:
        return seis
    n = int(np.prod(pshape))
    fd = {}
    for v, name in enumerate(var_names(c)):
        fd[name] = np.ascontiguousarray(fo[v * n:(v + 1) * n].reshape(pshape[::-1]).T)
    return seis, fd


def last_loop_seconds() -> float:
    """Wall-clock seconds of the time loop of the last forward call."""
    return float(load().or_last_loop_seconds())


def holberg(fdorder: int, maxrelerror: int):
    lib = load()
    hc = (ctypes.c_double * 7)()
    if lib.or_holberg(fdorder, maxrelerror, hc):
        raise ValueError("bad FDORDER/MAXRELERROR")
    return np.array(hc[:], dtype=np.float64)


def kernel_hc(fdorder: int, maxrelerror: int):
    lib = load()
    hc = np.zeros(7, np.float32)
    lib.or_kernel_hc(fdorder, maxrelerror, fp(hc))
    return hc


def cpml(npower, kmax, fpml, vppml, dh, dt, nab):
    lib = load()
    arrs = [np.zeros(2 * nab + 1, np.float32) for _ in range(6)]
    st = lib.or_cpml_coeff(npower, kmax, fpml, vppml, dh, dt, nab, *[fp(a) for a in arrs])
    K, b, a, Kh, bh, ah = [x[:2 * nab] for x in arrs]
    return dict(K=K, b=b, a=a, K_half=Kh, b_half=bh, a_half=ah, state=st)


def taper(nab, abpc):
    lib = load()
    t = np.zeros(nab, np.float32)
    lib.or_taper(nab, abpc, fp(t))
    return t


def transform(params: dict, **csts):
    """Reference parameter transforms; returns a dict of [NZ, (NY,) NX] arrays."""
    lib = load()
    c = make_csts(**csts)
    p0, p1, rho, taup, taus = _params(params, c)
    shape = [int(v) for v in np.atleast_1d(csts["N"])]
    n = int(np.prod(shape))
    names = ["M", "mu", "rho", "rip", "rjp", "rkp", "muipkp", "muipjp", "mujpkp", "taup",
             "taus", "tausipkp", "tausipjp", "tausjpkp"]
    outs = {k: np.zeros(n, np.float32) for k in names}
    lib.or_transform(ctypes.byref(c), fp(p0), fp(p1), fp(rho), fp(taup), fp(taus),
                     *[fp(outs[k]) for k in names])
    return {k: np.ascontiguousarray(v.reshape(shape[::-1]).T) for k, v in outs.items()}


def check_stability(params: dict, **csts):
    lib = load()
    c = make_csts(**csts)
    t = transform(params, **csts)
    dts = ctypes.c_float()
    M = to_c(t["M"])
    mu = to_c(t["mu"]) if c.ND != 22 else None
    rho = to_c(t["rho"])
    bad = lib.or_check_stability(ctypes.byref(c), fp(M), fp(mu), fp(rho), ctypes.byref(dts))
    return bool(bad), float(dts.value)


FIX_RES_SCALE, FIX_ADJ_CPML, FIX_RHO_SCALE, FIX_SURF_SYY, FIX_VISCO_F0 = 1, 2, 4, 8, 16
FIX_ALL = 223
FIX_RMSNORM = 64
FIX_HESS = 128
HISTORY = 32    # elastic runs also take the forward state from the stored history


def gradient(params: dict, src_pos, src, rec_pos, din=None, res=None, gradfix=FIX_ALL,
             nthreads: int = 0, fields: bool = False, movout: int = 0, **csts):
    """Oracle gradient of one shot by backpropagation.

    din: observed seismograms (list of (NT, nrec) in out_names order), or
    res: adjoint sources in the same layout (used as given, then scaled).
    Returns dict(out=[(NT,nrec)...], grad={p0name, p1name, 'rho'} arrays in
    model layout, res=[scaled adjoint sources], rms=(rms, rmsnorm)) and,
    with fields=True, the backpropagated forward wavefields at the end.
    movout > 0 adds movie={name: (NT/movout, *N) interior frames} of the
    gradient run (frames no adjoint step writes are zero)."""
    lib = load()
    c = make_csts(nthreads=nthreads, **csts)
    p0, p1, rho, taup, taus = _params(params, c)
    sp = np.ascontiguousarray(np.asarray(src_pos, np.float32).T)
    s = np.ascontiguousarray(np.asarray(src, np.float32).T)
    rp = np.ascontiguousarray(np.asarray(rec_pos, np.float32).T)
    nout = lib.or_nout(ctypes.byref(c))
    nrec = rp.shape[0]

    def pack(lst):
        if lst is None:
            return None
        a = np.stack([np.asarray(x, np.float32).T for x in lst])
        return np.ascontiguousarray(a.reshape(nout, nrec, c.NT))

    d = pack(din)
    r = pack(res)
    out = np.zeros((nout, nrec, c.NT), np.float32)
    res_out = np.zeros((nout, nrec, c.NT), np.float32)
    shape = [int(v) for v in np.atleast_1d(csts["N"])]
    n = int(np.prod(shape))
    grad = np.zeros(3 * n, np.float32)
    rms = (ctypes.c_double * 2)()
    H = c.FDORDER // 2
    pshape = [v + 2 * H for v in shape]
    nv = int(np.prod(pshape))
    nvars = lib.or_nvars(ctypes.byref(c))
    fo = np.zeros(nvars * nv, np.float32) if fields else None
    hess = np.zeros(3 * n, np.float32) if csts.get("Hout") else None
    gtau = np.zeros(2 * n, np.float32) if c.L > 0 and c.param_type <= 2 else None
    nfr = c.NT // movout if movout > 0 else 0
    mov = np.zeros(nvars * nfr * nv, np.float32) if nfr else None
    err = ctypes.create_string_buffer(256)
    st = lib.or_gradient(ctypes.byref(c), fp(p0), fp(p1), fp(rho), fp(taup), fp(taus),
                         sp.shape[0], fp(sp), fp(s), nrec, fp(rp), fp(d), fp(r),
                         int(gradfix), fp(out), fp(res_out), fp(grad), rms, fp(fo),
                         fp(hess), fp(gtau), int(movout), fp(mov), err, 256)
    if st:
        raise RuntimeError("oracle: " + err.value.decode())
    names = {0: ("vp", "vs"), 1: ("M", "mu"), 2: ("Ip", "Is"), 3: ("vp", "vs")}[c.param_type]
    g = {}
    for q, name in enumerate((names[0], names[1], "rho")):
        g[name] = np.ascontiguousarray(grad[q * n:(q + 1) * n].reshape(shape[::-1]).T)
    result = dict(out=[np.ascontiguousarray(o.T) for o in out], grad=g,
                  res=[np.ascontiguousarray(o.T) for o in res_out], rms=(rms[0], rms[1]))
    if gtau is not None:
        for q, name in enumerate(("taup", "taus")):
            g[name] = np.ascontiguousarray(gtau[q * n:(q + 1) * n].reshape(shape[::-1]).T)
    if hess is not None:
        result["hess"] = {name: np.ascontiguousarray(hess[q * n:(q + 1) * n].reshape(shape[::-1]).T)
                          for q, name in enumerate((names[0], names[1], "rho"))}
    if fields:
        fd = {}
        for v, name in enumerate(var_names(c)):
            fd[name] = np.ascontiguousarray(fo[v * nv:(v + 1) * nv].reshape(pshape[::-1]).T)
        result["fields"] = fd
    if mov is not None:
        inner = tuple(slice(H, -H) for _ in shape)
        mv = {}
        for v, name in enumerate(var_names(c)):
            fr = mov[v * nfr * nv:(v + 1) * nfr * nv].reshape([nfr] + pshape[::-1])
            fr = np.moveaxis(fr, 0, -1).T            # (nfr, *pshape) with z first
            mv[name] = np.ascontiguousarray(fr[(slice(None),) + inner])
        result["movie"] = mv
    return result


def shell_check(**csts):
    """(inconsistencies, NBND) of the backpropagation boundary shell."""
    lib = load()
    c = make_csts(**csts)
    n = ctypes.c_long()
    bad = lib.or_shell_check(ctypes.byref(c), ctypes.byref(n))
    return int(bad), int(n.value)


def dft_sampling(freqs, **csts):
    """(DTNYQ, NTNYQ, df, bins) of the DFT gradient (or_dft_sampling)."""
    lib = load()
    c = make_csts(**csts)
    f = np.ascontiguousarray(np.asarray(freqs, np.float32).ravel())
    bins = np.zeros(f.size, np.float32)
    dt, nt, df = ctypes.c_int(), ctypes.c_int(), np.zeros(1, np.float32)
    st = lib.or_dft_sampling(ctypes.byref(c), f.size, fp(f), ctypes.byref(dt), ctypes.byref(nt),
                             fp(df), fp(bins))
    if st == 1:
        raise RuntimeError("oracle: bad gradfreqs")
    return dt.value, nt.value, float(df[0]), bins


def dft_coef(pt, visco, M, mu, taup, taus, rho, ND, L, al):
    """The 24 scalar-product coefficients of one cell (or_dft_coef)."""
    c = (ctypes.c_double * 24)()
    load().or_dft_coef(c, int(pt), int(visco), M, mu, taup, taus, rho, ND, L, al)
    return np.array(c[:])


def gradient_dft(params: dict, src_pos, src, rec_pos, freqs, din=None, res=None,
                 gradfix=FIX_ALL, nthreads: int = 0, spectra: bool = False, **csts):
    """Oracle DFT-domain gradient of one shot (back_prop_type 2).

    Returns dict(out=[(NT,nrec)...], grad={p0name, p1name, 'rho'[, 'taup',
    'taus']} in model layout (double), rms=(rms, rmsnorm) at the gradient
    frequencies) and, with spectra=True, spectra: complex array
    [nq, nf, *padded N] of the forward DFTs (or_gradient_dft)."""
    lib = load()
    c = make_csts(nthreads=nthreads, **csts)
    p0, p1, rho, taup, taus = _params(params, c)
    sp = np.ascontiguousarray(np.asarray(src_pos, np.float32).T)
    s = np.ascontiguousarray(np.asarray(src, np.float32).T)
    rp = np.ascontiguousarray(np.asarray(rec_pos, np.float32).T)
    nout = lib.or_nout(ctypes.byref(c))
    nrec = rp.shape[0]

    def pack(lst):
        if lst is None:
            return None
        a = np.stack([np.asarray(x, np.float32).T for x in lst])
        return np.ascontiguousarray(a.reshape(nout, nrec, c.NT))

    f = np.ascontiguousarray(np.asarray(freqs, np.float32).ravel())
    out = np.zeros((nout, nrec, c.NT), np.float32)
    shape = [int(v) for v in np.atleast_1d(csts["N"])]
    n = int(np.prod(shape))
    grad = np.zeros(5 * n, np.float64)
    rms = (ctypes.c_double * 2)()
    H = c.FDORDER // 2
    pshape = [v + 2 * H for v in shape]
    nv = int(np.prod(pshape))
    nq = (9 + 6 * c.L) if c.ND == 3 else (5 + 3 * c.L)
    so = np.zeros(nq * f.size * nv * 2, np.float32) if spectra else None
    err = ctypes.create_string_buffer(256)
    st = lib.or_gradient_dft(ctypes.byref(c), fp(p0), fp(p1), fp(rho), fp(taup), fp(taus),
                             sp.shape[0], fp(sp), fp(s), nrec, fp(rp), fp(pack(din)),
                             fp(pack(res)), int(gradfix), f.size, fp(f), fp(out),
                             grad.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), rms,
                             fp(so), err, 256)
    if st:
        raise RuntimeError("oracle: " + err.value.decode())
    names = {0: ("vp", "vs"), 1: ("M", "mu"), 2: ("Ip", "Is"), 3: ("vp", "vs")}[c.param_type]
    keys = [names[0], names[1], "rho"] + (["taup", "taus"] if c.L > 0 else [])
    g = {k: np.ascontiguousarray(grad[q * n:(q + 1) * n].reshape(shape[::-1]).T)
         for q, k in enumerate(keys)}
    result = dict(out=[np.ascontiguousarray(o.T) for o in out], grad=g, rms=(rms[0], rms[1]))
    if spectra:
        z = so.reshape(nq, f.size, nv, 2)
        z = z[..., 0] + 1j * z[..., 1]
        result["spectra"] = np.moveaxis(z.reshape([nq, f.size] + pshape[::-1]),
                                        list(range(2, 2 + len(shape))),
                                        list(range(2 + len(shape) - 1, 1, -1)))
    return result
