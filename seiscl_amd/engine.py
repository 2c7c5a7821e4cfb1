"""Python handles over the C-ABI: a GPU context and a resident model.

Array conventions are SeisCL.py's (SeisCL/SeisCL.py:414-602):
  - model parameters are numpy arrays shaped [NZ, NX] or [NZ, NY, NX];
  - src_pos is (5, nsrc), src is (NT, nsrc), rec_pos is (8, nrec);
  - seismograms come back as (NT, nrec) arrays, one per output variable.
The C side receives the transposed (x-slowest, z-fastest) buffers, exactly
what the reference's HDF5 files hold after SeisCL.py's np.transpose.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import SeisCsts, check, fptr

# SeisCL.py constructor defaults (SeisCL/SeisCL.py:36-83) for the keys the
# engine uses.
CSTS_DEFAULTS = dict(
    ND=2, dh=10.0, dt=0.0008, NT=875, L=0, f0=15.0, FL=(15.0,), FDORDER=8,
    MAXRELERROR=1, FP16=0, freesurf=0, abs_type=1, VPPML=3500.0, NPOWER=2.0,
    FPML=15.0, K_MAX_CPML=2.0, nab=16, abpc=6.0, seisout=2, param_type=0,
    tmax=0.0, tmin=0.0, gradout=0, gradfix=223, restype=0, scalerms=0, scaleshot=0,
    scalermsnorm=0, fmin=0.0, fmax=0.0, Hout=0,
)

PARAM_NAMES = {0: ("vp", "vs", "rho"), 1: ("M", "mu", "rho"), 2: ("Ip", "Is", "rho")}


def make_csts(**kw) -> SeisCsts:
    """Build a seis_csts from SeisCL.py-style keyword constants."""
    d = dict(CSTS_DEFAULTS)
    d.update(kw)
    c = SeisCsts()
    _lib.load().seis_csts_default(ctypes.byref(c))
    N = [int(v) for v in np.atleast_1d(d["N"])]
    c.ND = int(d["ND"])
    for i, v in enumerate(N):
        c.N[i] = v
    c.NT = int(d["NT"])
    c.dt = float(d["dt"])
    c.dh = float(d["dh"])
    c.FDORDER = int(d["FDORDER"])
    c.MAXRELERROR = int(d["MAXRELERROR"])
    c.freesurf = int(d["freesurf"])
    c.abs_type = int(d["abs_type"])
    c.nab = int(d["nab"])
    c.abpc = float(d["abpc"])
    c.VPPML = float(d["VPPML"])
    c.NPOWER = float(d["NPOWER"])
    c.FPML = float(d["FPML"])
    c.K_MAX_CPML = float(d["K_MAX_CPML"])
    c.L = int(d["L"])
    c.f0 = float(d["f0"])
    fl = np.atleast_1d(np.asarray(d["FL"], dtype=np.float32))
    for i in range(min(len(fl), _lib.MAXL)):
        c.FL[i] = float(fl[i])
    c.seisout = int(d["seisout"])
    c.param_type = int(d["param_type"])
    # tmax/tmin are seconds in SeisCL.py; the C side counts steps
    # (src/read_hdf5.c:523-528)
    c.tmax = int(np.float32(d["tmax"]) / np.float32(c.dt)) if d["tmax"] else 0
    c.tmin = int(np.float32(d["tmin"]) / np.float32(c.dt)) if d["tmin"] else 0
    c.gradout = int(d["gradout"])
    c.FP16 = int(d["FP16"])
    c.gradfix = int(d["gradfix"])
    c.restype = int(d["restype"])
    c.scalerms = int(d["scalerms"])
    c.scaleshot = int(d["scaleshot"])
    c.scalermsnorm = int(d["scalermsnorm"])
    c.fmin = float(d["fmin"])
    c.fmax = float(d["fmax"])
    c.Hout = int(d["Hout"])
    return c


def _pack_traces(lst) -> np.ndarray:
    """list of (NT, nrec) arrays -> C array [nout][nrec][NT]."""
    return np.ascontiguousarray(np.stack([np.asarray(x, dtype=np.float32).T for x in lst]))


def misfit(out, din, **csts):
    """Residuals of a shot (var_res_raw / rtm_res with the misfit options in
    csts: restype, scalerms, scaleshot, scalermsnorm, fmin, fmax): host
    code of the library, no device needed.  out, din: lists of (NT, nrec)
    arrays, one per output; returns (res list, (rms, rmsnorm))."""
    c = make_csts(**csts)
    o, d = _pack_traces(out), _pack_traces(din)
    r = np.zeros_like(o)
    rms, rn = ctypes.c_double(), ctypes.c_double()
    check(_lib.load().seis_misfit(ctypes.byref(c), o.shape[0], o.shape[1], fptr(o), fptr(d),
                                  fptr(r), ctypes.byref(rms), ctypes.byref(rn)))
    return [np.ascontiguousarray(x.T) for x in r], (rms.value, rn.value)


def to_c_layout(a: np.ndarray) -> np.ndarray:
    """[NZ, (NY,) NX] numpy array -> flat float32, x slowest, z fastest."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).T).ravel()


def from_c_layout(flat: np.ndarray, shape) -> np.ndarray:
    """Inverse of to_c_layout for a grid of the given [NZ, (NY,) NX] shape."""
    return np.ascontiguousarray(flat.reshape(tuple(shape)[::-1]).T)


def param_arrays(params: dict, param_type: int, L: int, ND: int):
    """Order the parameter dict as the C-ABI wants it (p0, p1, rho, taup, taus)."""
    names = PARAM_NAMES.get(param_type)
    if names is None:
        names = ("vpR", "vsR", "rho")
    out = []
    for k, n in enumerate(names):
        if ND == 22 and k == 1:
            out.append(None)
            continue
        if param_type == 3 and k < 2:
            n = ("vpR", "vsR")[k]
        out.append(to_c_layout(params[n]))
    if L > 0:
        tp = "vpI" if param_type == 3 else "taup"
        ts = "vsI" if param_type == 3 else "taus"
        out.append(to_c_layout(params[tp]))
        out.append(to_c_layout(params[ts]) if ND != 22 else None)
    else:
        out += [None, None]
    return out


def shot_arrays(src_pos, src, rec_pos):
    """SeisCL.py (5,ns)/(NT,ns)/(8,nr) arrays -> C row-major float32 buffers."""
    sp = np.ascontiguousarray(np.asarray(src_pos, dtype=np.float32).T)
    s = np.ascontiguousarray(np.asarray(src, dtype=np.float32).T)
    rp = np.ascontiguousarray(np.asarray(rec_pos, dtype=np.float32).T)
    return sp, s, rp


class Context:
    """One GPU (seis_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        check(self.lib.seis_ctx_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = int(device)

    def footprint(self) -> int:
        return int(self.lib.seis_ctx_footprint(self.handle))

    def sync(self):
        check(self.lib.seis_ctx_sync(self.handle))

    def close(self):
        if self.handle:
            self.lib.seis_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Model:
    """A modeling case resident on one GPU (seis_model)."""

    def __init__(self, ctx: Context, params: dict, **csts):
        self.ctx = ctx
        self.lib = ctx.lib
        self.csts = make_csts(**csts)
        self.shape = tuple(int(v) for v in np.atleast_1d(csts["N"]))
        self._keep = param_arrays(params, self.csts.param_type, self.csts.L, self.csts.ND)
        h = ctypes.c_void_p()
        check(self.lib.seis_model_create(ctx.handle, ctypes.byref(self.csts),
                                         *[fptr(a) for a in self._keep], ctypes.byref(h)))
        self.handle = h
        n = self.lib.seis_out_count(h)
        self.out_names = [self.lib.seis_out_name(h, i).decode() for i in range(n)]

    def set_params(self, params: dict):
        self._keep = param_arrays(params, self.csts.param_type, self.csts.L, self.csts.ND)
        check(self.lib.seis_model_set_params(self.handle, *[fptr(a) for a in self._keep]))

    def footprint(self) -> int:
        return int(self.lib.seis_model_footprint(self.handle))

    def set_profiling(self, on: bool):
        check(self.lib.seis_model_set_profiling(self.handle, int(bool(on))))

    def timing(self):
        v, s, loop = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(self.lib.seis_last_timing(self.handle, ctypes.byref(v), ctypes.byref(s),
                                        ctypes.byref(loop)))
        return dict(update_v_ms=v.value, update_s_ms=s.value, loop_ms=loop.value)

    def forward(self, src_pos, src, rec_pos, raw: bool = False):
        """Model one shot; returns a list of (NT, nrec) seismograms in
        ``self.out_names`` order (raw=True: the C array [nout][nrec][NT])."""
        sp, s, rp = shot_arrays(src_pos, src, rec_pos)
        nsrc, nrec = sp.shape[0], rp.shape[0]
        NT = self.csts.NT
        out = np.zeros((len(self.out_names), nrec, NT), dtype=np.float32)
        check(self.lib.seis_forward_shot(self.handle, nsrc, fptr(sp), fptr(s), nrec,
                                         fptr(rp), fptr(out)))
        if raw:
            return out
        return [np.ascontiguousarray(o.T) for o in out]

    def bind_shot(self, src_pos, src, rec_pos, init: bool = True):
        """Upload one shot and (init) zero the wavefields (first half of forward)."""
        self._shot = shot_arrays(src_pos, src, rec_pos)
        sp, s, rp = self._shot
        check(self.lib.seis_shot_bind(self.handle, sp.shape[0], fptr(sp), fptr(s), rp.shape[0],
                                      fptr(rp)))
        if init:
            check(self.lib.seis_launch_init(self.handle))

    def run_steps(self, t0: int, nsteps: int):
        """Advance the bound shot by nsteps time steps starting at step t0."""
        check(self.lib.seis_run_steps(self.handle, int(t0), int(nsteps)))

    def read_shot(self, f64: bool = False):
        """Seismograms of the bound shot, (NT, nrec) per output; f64: in
        double, for models running in FP64 (FP16=64)."""
        nrec = self._shot[2].shape[0]
        dt = np.float64 if f64 else np.float32
        out = np.zeros((len(self.out_names), nrec, self.csts.NT), dtype=dt)
        if f64:
            check(self.lib.seis_shot_read_f64(
                self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        else:
            check(self.lib.seis_shot_read(self.handle, fptr(out)))
        return [np.ascontiguousarray(o.T) for o in out]

    def read_field(self, name: str) -> np.ndarray:
        """Wavefield in the reference's padded layout, reshaped [NZp, (NYp,) NXp]."""
        H = self.csts.FDORDER // 2
        shp = [n + 2 * H for n in self.shape]
        flat = np.zeros(int(np.prod(shp)), dtype=np.float32)
        check(self.lib.seis_read_field(self.handle, name.encode(), fptr(flat)))
        return from_c_layout(flat, shp)

    # ---- gradient by backpropagation --------------------------------------
    def _pack(self, lst, nrec):
        """list of (NT, nrec) arrays in out_names order -> [nout][nrec][NT]."""
        a = np.stack([np.asarray(x, dtype=np.float32).T for x in lst])
        return np.ascontiguousarray(a.reshape(len(self.out_names), nrec, self.csts.NT))

    def grad_zero(self):
        check(self.lib.seis_grad_zero(self.handle))

    def set_grad_memory(self, nbytes: int):
        """Budget of device bytes for the viscoelastic gradient's checkpoints
        and recomputed history (0: 90% of the free device memory)."""
        check(self.lib.seis_grad_set_memory(self.handle, int(nbytes)))

    def grad_plan(self) -> dict:
        """Checkpoint plan of the last viscoelastic forward_save."""
        lv, n, n0 = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        K = (ctypes.c_int * 4)()
        b = ctypes.c_size_t()
        check(self.lib.seis_grad_plan(self.handle, ctypes.byref(lv), K, ctypes.byref(n),
                                      ctypes.byref(n0), ctypes.byref(b)))
        return dict(levels=lv.value, K=list(K)[:lv.value], n=n.value, n0=n0.value,
                    bytes=b.value)

    def forward_save(self, src_pos, src, rec_pos):
        """Forward run that keeps the boundary shell of every step for the
        adjoint run; returns the seismograms like forward()."""
        self._shot = shot_arrays(src_pos, src, rec_pos)
        sp, s, rp = self._shot
        out = np.zeros((len(self.out_names), rp.shape[0], self.csts.NT), dtype=np.float32)
        check(self.lib.seis_forward_save(self.handle, sp.shape[0], fptr(sp), fptr(s),
                                         rp.shape[0], fptr(rp), fptr(out)))
        return [np.ascontiguousarray(o.T) for o in out]

    def saved_read(self) -> np.ndarray:
        """The saved forward run of the last forward_save as one host array
        (checkpoint_d2h of src/time_stepping.c:92-168): what seis_adjoint
        needs, so that a later process can run the adjoint without the
        forward run."""
        n = ctypes.c_size_t()
        check(self.lib.seis_grad_saved_size(self.handle, ctypes.byref(n)))
        buf = np.zeros(n.value // 4, dtype=np.float32)
        check(self.lib.seis_grad_saved_read(self.handle, fptr(buf)))
        return buf

    def saved_write(self, src_pos, src, rec_pos, saved: np.ndarray):
        """Bind the shot and restore its saved forward run (checkpoint_h2d,
        src/time_stepping.c:170-270); adjoint() then runs without a forward
        run."""
        self.bind_shot(src_pos, src, rec_pos, init=False)
        buf = np.ascontiguousarray(saved, dtype=np.float32)
        check(self.lib.seis_grad_saved_write(self.handle, fptr(buf)))

    def adjoint(self, res, scale: bool = True, movout: int = 0, movie_vars=None):
        """Adjoint run of the shot of the last forward_save with adjoint
        sources res (list of (NT, nrec)); res_scale is applied when scale.

        movout > 0 records the movie of the gradient run (seis_adjoint_movie)
        for the wavefields movie_vars and returns {name: array [NZ, (NY,) NX,
        NT // movout]}: frame (t+1)//movout - 1 holds the backpropagated
        forward wavefield after the adjoint step t, t % movout == 0."""
        rp = self._shot[2]
        r = self._pack(res, rp.shape[0])
        if scale:
            check(self.lib.seis_res_scale(self.handle, rp.shape[0], fptr(rp), fptr(r)))
        if movout <= 0:
            check(self.lib.seis_adjoint(self.handle, fptr(r)))
            return None
        names = list(movie_vars or [])
        nfr = int(self.csts.NT) // int(movout)
        n = int(np.prod(self.shape))
        frames = np.zeros(len(names) * max(nfr, 1) * n, dtype=np.float32)
        cn = (ctypes.c_char_p * max(len(names), 1))(*[x.encode() for x in names])
        check(self.lib.seis_adjoint_movie(self.handle, fptr(r), int(movout), len(names), cn,
                                          fptr(frames)))
        out = {}
        for v, name in enumerate(names):
            fr = [from_c_layout(frames[(v * nfr + f) * n:(v * nfr + f + 1) * n], self.shape)
                  for f in range(nfr)]
            out[name] = np.stack(fr, axis=-1) if fr else np.zeros(tuple(self.shape) + (0,),
                                                                  np.float32)
        return out

    def res_scale(self, res, rec_pos):
        """res_scale (src/residuals.c:452-570) of adjoint sources res (list of
        (NT, nrec)) for receivers rec_pos (8, nrec); returns the scaled list."""
        rp = np.ascontiguousarray(np.asarray(rec_pos, np.float32).T)
        r = self._pack(res, rp.shape[0])
        check(self.lib.seis_res_scale(self.handle, rp.shape[0], fptr(rp), fptr(r)))
        return [np.ascontiguousarray(x.T) for x in r]

    def residuals(self, out, din):
        """Residuals of the bound shot's outputs against din with the model's
        misfit options (seis_residuals); returns (res list, (rms, rmsnorm))."""
        o, d = _pack_traces(out), _pack_traces(din)
        r = np.zeros_like(o)
        rms, rn = ctypes.c_double(), ctypes.c_double()
        check(self.lib.seis_residuals(self.handle, o.shape[1], fptr(o), fptr(d), fptr(r),
                                      ctypes.byref(rms), ctypes.byref(rn)))
        return [np.ascontiguousarray(x.T) for x in r], (rms.value, rn.value)

    def grad_read(self) -> dict:
        """Accumulated gradients {p0, p1, rho} in the model layout, plus
        {taup, taus} (at fixed wave speeds) for viscoelastic models."""
        n = int(np.prod(self.shape))
        g = np.zeros(3 * n, dtype=np.float32)
        check(self.lib.seis_grad_read(self.handle, fptr(g)))
        names = PARAM_NAMES.get(self.csts.param_type, ("vp", "vs", "rho"))
        out = {nm: from_c_layout(g[q * n:(q + 1) * n], self.shape) for q, nm in enumerate(names)}
        if self.csts.L > 0 and self.csts.param_type <= 2:
            t = np.zeros(2 * n, dtype=np.float32)
            check(self.lib.seis_grad_read_tau(self.handle, fptr(t)))
            out["taup"] = from_c_layout(t[:n], self.shape)
            out["taus"] = from_c_layout(t[n:], self.shape)
        return out

    # ---- device-resident variants (torch tensors on the model's GPU) -----

    def _c_layout_device(self, t):
        """A torch tensor [NZ, (NY,) NX] -> float32, C layout (z fastest),
        contiguous, on its device."""
        import torch
        t = t.detach().to(torch.float32)
        return t.permute(*reversed(range(t.dim()))).contiguous()

    def set_params_device(self, params: dict):
        """set_params from torch tensors on the GPU (names as set_params)."""
        import torch
        names = PARAM_NAMES.get(self.csts.param_type, ("vpR", "vsR", "rho"))
        if self.csts.param_type == 3:
            names = ("vpR", "vsR", "rho")
        keep = []
        for k, n in enumerate(names):
            if self.csts.ND == 22 and k == 1:
                keep.append(None)
                continue
            keep.append(self._c_layout_device(params[n]))
        if self.csts.L > 0:
            tp = "vpI" if self.csts.param_type == 3 else "taup"
            ts = "vsI" if self.csts.param_type == 3 else "taus"
            keep.append(self._c_layout_device(params[tp]))
            keep.append(self._c_layout_device(params[ts]) if self.csts.ND != 22 else None)
        else:
            keep += [None, None]
        torch.cuda.current_stream().synchronize()
        ptr = [ctypes.c_void_p(t.data_ptr()) if t is not None else None for t in keep]
        check(self.lib.seis_model_set_params_device(self.handle, *ptr))

    def read_shot_device(self):
        """Seismograms of the bound shot as a torch tensor [nout, nrec, NT]
        on the GPU (no host round trip)."""
        import torch
        nrec = self._shot[2].shape[0]
        out = torch.empty((len(self.out_names), nrec, self.csts.NT), dtype=torch.float32,
                          device="cuda")
        torch.cuda.current_stream().synchronize()
        check(self.lib.seis_shot_read_device(self.handle, ctypes.c_void_p(out.data_ptr())))
        return out

    def forward_save_device(self, src_pos, src, rec_pos):
        """forward_save whose seismograms stay on the GPU: [nout, nrec, NT]."""
        self._shot = shot_arrays(src_pos, src, rec_pos)
        sp, s, rp = self._shot
        check(self.lib.seis_forward_save(self.handle, sp.shape[0], fptr(sp), fptr(s),
                                         rp.shape[0], fptr(rp), None))
        return self.read_shot_device()

    def adjoint_device(self, res, scale: bool = True):
        """adjoint() with the adjoint sources a torch tensor [nout, nrec, NT]
        on the GPU; res_scale runs on the device when scale."""
        import torch
        r = res.detach().to(torch.float32).contiguous()
        torch.cuda.current_stream().synchronize()
        check(self.lib.seis_adjoint_device(self.handle, ctypes.c_void_p(r.data_ptr()),
                                           int(bool(scale))))

    def grad_read_device(self) -> dict:
        """grad_read() as torch tensors on the GPU, in the model layout."""
        import torch
        n = int(np.prod(self.shape))
        g = torch.empty(3 * n, dtype=torch.float32, device="cuda")
        visco = self.csts.L > 0 and self.csts.param_type <= 2
        t = torch.empty(2 * n, dtype=torch.float32, device="cuda") if visco else None
        torch.cuda.current_stream().synchronize()
        check(self.lib.seis_grad_read_device(
            self.handle, ctypes.c_void_p(g.data_ptr()),
            ctypes.c_void_p(t.data_ptr()) if visco else None))
        rev = tuple(self.shape)[::-1]

        def model_layout(flat):
            return flat.reshape(rev).permute(*reversed(range(len(rev)))).contiguous()
        names = PARAM_NAMES.get(self.csts.param_type, ("vp", "vs", "rho"))
        out = {nm: model_layout(g[q * n:(q + 1) * n]) for q, nm in enumerate(names)}
        if visco:
            out["taup"] = model_layout(t[:n])
            out["taus"] = model_layout(t[n:])
        return out

    def hess_read(self) -> dict:
        """Approximate Hessian {p0, p1, rho} accumulated with the gradient
        (Hout=1), transformed like the gradient."""
        n = int(np.prod(self.shape))
        h = np.zeros(3 * n, dtype=np.float32)
        check(self.lib.seis_hess_read(self.handle, fptr(h)))
        names = PARAM_NAMES.get(self.csts.param_type, ("vp", "vs", "rho"))
        return {nm: from_c_layout(h[q * n:(q + 1) * n], self.shape) for q, nm in enumerate(names)}

    def grad_dft(self, freqs=None):
        """Select the DFT-domain gradient (back_prop_type 2) at the given
        frequencies (Hz) for the next gradient runs; None or an empty list
        returns to backpropagation (seis_grad_dft)."""
        f = np.ascontiguousarray(np.asarray([] if freqs is None else freqs, np.float32).ravel())
        check(self.lib.seis_grad_dft(self.handle, int(f.size), fptr(f) if f.size else None))

    def gradient(self, src_pos, src, rec_pos, din=None, res=None):
        """One shot of gradient (added to the accumulators): forward with
        boundary saving, residuals against din (or adjoint sources res),
        res_scale, adjoint.  Returns (seismograms, (rms, rmsnorm))."""
        sp, s, rp = shot_arrays(src_pos, src, rec_pos)
        self._shot = (sp, s, rp)
        nrec = rp.shape[0]
        d = self._pack(din, nrec) if din is not None else None
        r = self._pack(res, nrec) if res is not None else None
        out = np.zeros((len(self.out_names), nrec, self.csts.NT), dtype=np.float32)
        rms = (ctypes.c_double * 2)()
        check(self.lib.seis_gradient_shot(self.handle, sp.shape[0], fptr(sp), fptr(s), nrec,
                                          fptr(rp), fptr(d), fptr(r), fptr(out), rms))
        return [np.ascontiguousarray(o.T) for o in out], (rms[0], rms[1])

    def close(self):
        if getattr(self, "handle", None):
            self.lib.seis_model_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Modeling jobs stepped together (seis_batch): one launch per kernel
    class and time step advances every job in the batch."""

    def __init__(self, ctx: Context, max_slots: int = 64):
        self.ctx = ctx
        self.lib = ctx.lib
        h = ctypes.c_void_p()
        check(self.lib.seis_batch_create(ctx.handle, int(max_slots), ctypes.byref(h)))
        self.handle = h
        self.max_slots = int(max_slots)

    def add(self, model: Model) -> int:
        """Add a model whose shot is bound (Model.bind_shot); returns its slot."""
        slot = ctypes.c_int()
        check(self.lib.seis_batch_add(self.handle, model.handle, ctypes.byref(slot)))
        return slot.value

    def remove(self, slot: int):
        check(self.lib.seis_batch_remove(self.handle, int(slot)))

    def step(self, nsteps: int):
        check(self.lib.seis_batch_step(self.handle, int(nsteps)))

    def status(self):
        """(jobs still stepping, fewest steps any of them has left)"""
        n, m = ctypes.c_int(), ctypes.c_int()
        check(self.lib.seis_batch_status(self.handle, ctypes.byref(n), ctypes.byref(m)))
        return n.value, m.value

    def left(self, slot: int) -> int:
        return int(self.lib.seis_batch_slot_left(self.handle, int(slot)))

    def close(self):
        if getattr(self, "handle", None):
            self.lib.seis_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
