"""In-process drop-in for the SeisCL Python class (SeisCL/SeisCL.py:33-1052).

``SeisCL`` keeps the reference class's constructor keywords, attribute
names, acquisition helpers and the call sequence of an FWI script:

    seis = SeisCL(N=..., ND=2, ...); seis.surface_acquisition_2d()
    seis.set_forward(srcids, params, withgrad=True)
    seis.execute()
    data = seis.read_data(); grad = seis.read_grad(); rms = seis.read_rms()

Two ways to run:

- in process (default): ``execute`` runs the shots on the GPU through the
  C-ABI and keeps the outputs in memory, so the read_* methods return what
  the reference reads back from its output files (same names, shapes and
  trace order); file paths are accepted and ignored.  Shots are the
  distinct srcids of the selected sources; receivers belong to the shot
  whose srcid they carry (SeisCL.py:765-808).
- ``files=True``: the reference's workflow (SeisCL.py:414-505, 507-763).
  set_forward / set_backward / write_data write the HDF5 constants, model,
  residual and data files; ``execute`` runs the SeisCL_MPI program on them
  (a child process, or its body in this process with ``spawn=False``); the
  read_* methods read the output files.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np

from . import h5mat
from .engine import Context, Model

# keys of the constants file (SeisCL/SeisCL.py:23-30)
CSTS_KEYS = ['N', 'ND', 'dh', 'dt', 'NT', 'freesurf', 'FDORDER', 'MAXRELERROR', 'L', 'f0',
             'FL', 'src_pos', 'rec_pos', 'src', 'abs_type', 'VPPML', 'NPOWER', 'FPML',
             'K_MAX_CPML', 'nab', 'abpc', 'pref_device_type', 'no_use_GPUs', 'MPI_NPROC_SHOT',
             'nmax_dev', 'back_prop_type', 'param_type', 'gradfreqs', 'tmax', 'tmin',
             'scalerms', 'scalermsnorm', 'scaleshot', 'fmin', 'fmax', 'gradout', 'Hout',
             'gradsrcout', 'seisout', 'resout', 'rmsout', 'movout', 'restype', 'inputres',
             'FP16']


class SeisCLError(Exception):
    pass


class SeisCL:
    # constants passed to the engine (the reference's csts list, less the
    # parallelisation and file keys)
    _ENGINE_KEYS = ("ND", "N", "dh", "dt", "NT", "L", "f0", "FL", "FDORDER", "MAXRELERROR",
                    "FP16", "freesurf", "abs_type", "VPPML", "NPOWER", "FPML", "K_MAX_CPML",
                    "nab", "abpc", "seisout", "param_type", "tmax", "tmin", "restype",
                    "scalerms", "scaleshot", "scalermsnorm", "fmin", "fmax", "Hout")

    def __init__(self, N=None, ND=2, dh=10, dt=0.0008, NT=875, L=0, f0=15,
                 FL=np.array(15), FDORDER=8, MAXRELERROR=1, FP16=0,
                 src_pos_all=np.empty((5, 0)), rec_pos_all=np.empty((8, 0)), src_all=None,
                 freesurf=0, abs_type=1, VPPML=3500, NPOWER=2, FPML=15, K_MAX_CPML=2,
                 nab=16, abpc=6, with_docker=False, with_mpi=False, NP=1,
                 pref_device_type=4, MPI_NPROC_SHOT=1, nmax_dev=1,
                 no_use_GPUs=np.array([-1]), gradout=0, Hout=0, gradsrcout=0,
                 back_prop_type=1, cropgrad=True, gradfreqs=np.empty((1, 0)), param_type=0,
                 tmax=0, tmin=0, fmin=0, fmax=0, filter_offset=False,
                 offmin=-float("Inf"), offmax=float("Inf"), inputres=0, restype=0,
                 scalerms=0, scalermsnorm=0, scaleshot=0, seisout=2, resout=0, rmsout=0,
                 movout=0, file="SeisCL", workdir="./seiscl", device: int = 0,
                 gradfix: int = 223, files: bool = False, spawn: bool = True,
                 progname: str | None = None):
        self.N, self.ND, self.dh, self.dt, self.NT = N, ND, dh, dt, NT
        self.L, self.f0, self.FL = L, f0, FL
        self.FDORDER, self.MAXRELERROR, self.FP16 = FDORDER, MAXRELERROR, FP16
        self.src_pos_all, self.rec_pos_all, self.src_all = src_pos_all, rec_pos_all, src_all
        self.src_pos = self.rec_pos = self.src = None
        self.freesurf, self.abs_type = freesurf, abs_type
        self.VPPML, self.NPOWER, self.FPML, self.K_MAX_CPML = VPPML, NPOWER, FPML, K_MAX_CPML
        self.nab, self.abpc = nab, abpc
        self.with_docker, self.with_mpi, self.NP = with_docker, with_mpi, NP
        self.pref_device_type, self.MPI_NPROC_SHOT = pref_device_type, MPI_NPROC_SHOT
        self.nmax_dev, self.no_use_GPUs = nmax_dev, no_use_GPUs
        self.gradout, self.Hout, self.gradsrcout = gradout, Hout, gradsrcout
        self.back_prop_type, self.cropgrad, self.gradfreqs = back_prop_type, cropgrad, gradfreqs
        self.param_type, self.tmax, self.tmin = param_type, tmax, tmin
        self.fmin, self.fmax = fmin, fmax
        self.filter_offset, self.offmin, self.offmax = filter_offset, offmin, offmax
        self.inputres, self.restype = inputres, restype
        self.scalerms, self.scalermsnorm, self.scaleshot = scalerms, scalermsnorm, scaleshot
        self.seisout, self.resout, self.rmsout, self.movout = seisout, resout, rmsout, movout
        self.file, self.workdir = file, workdir
        self.device, self.gradfix = device, gradfix
        self.files, self.spawn = files, spawn
        self.progname = h5mat.BIN_PATH if progname is None else progname
        self.wavelet_generator = self.ricker_wavelet
        self._to_load_names = None
        self._params = None
        self._jobids = None
        self._din = None
        self._res_in = None
        self._results = {}
        self._ctx = None
        # saved forward runs of an inputres pass 1 run in process, by shot id
        self._saved = {}

    # ---- file names (SeisCL.py:355-371) ------------------------------------
    @property
    def file(self):
        return self._file

    @file.setter
    def file(self, file):
        self._file = file
        self.file_model = file + "_model.mat"
        self.file_csts = file + "_csts.mat"
        self.file_dout = file + "_dout.mat"
        self.file_gout = file + "_gout.mat"
        self.file_rms = file + "_rms.mat"
        self.file_movout = file + "_movie.mat"
        self.file_din = file + "_din.mat"
        self.file_res = file + "_res.mat"

    def _path(self, workdir, name):
        return os.path.join(self.workdir if workdir is None else workdir, name)

    # ---- names and shapes (SeisCL.py:341-412) -----------------------------
    @property
    def params(self):
        names = {0: ["vp", "vs", "rho"], 1: ["M", "mu", "rho"], 2: ["Ip", "Is", "rho"]}
        if self.param_type not in names:
            raise NotImplementedError()
        p = list(names[self.param_type])
        if self.ND == 22:
            p.remove(p[1])
        if self.L > 0:
            p += ["taup"] + (["taus"] if self.ND != 22 else [])
        return p

    @property
    def to_load_names(self):
        if self._to_load_names is not None:
            return self._to_load_names
        if self.seisout == 1:
            return {2: ["vx", "vz"], 22: ["vx", "vz"], 3: ["vx", "vy", "vz"]}[self.ND]
        if self.seisout == 2:
            return ["p"]
        if self.seisout == 3:
            return {2: ["sxx", "szz", "sxz"],
                    3: ["sxx", "syy", "szz", "sxz", "sxy", "syz"]}[self.ND]
        if self.seisout == 4:
            return {2: ["vx", "vz", "sxx", "szz", "sxz"], 22: ["vx", "vz", "p"],
                    3: ["vx", "vy", "vz", "sxx", "syy", "szz", "sxz", "sxy", "syz"]}[self.ND]
        return None

    @to_load_names.setter
    def to_load_names(self, values):
        self._to_load_names = values

    # ---- acquisition (SeisCL.py:765-952) ----------------------------------
    def prepare_data(self, srcids):
        if np.issubdtype(type(srcids), np.integer):
            srcids = [srcids]
        srcids = [int(s) for s in np.atleast_1d(srcids)]
        srcels = [i for i, el in enumerate(self.src_pos_all[3, :].astype(int)) if el in srcids]
        recels = [g for g, s in enumerate(self.rec_pos_all[3, :].astype(int)) if s in srcids]
        if len(srcels) <= 0:
            raise ValueError("No shot found")
        self.src_pos = self.src_pos_all[:, srcels]
        if self.src_all is None:
            self.src_all = np.stack([self.wavelet_generator()] * self.src_pos_all.shape[1], 1)
        self.src = self.src_all[:, srcels]
        self.rec_pos = self.rec_pos_all[:, recels]
        if self.filter_offset:
            valid = []
            for ii in range(self.rec_pos.shape[1]):
                sid = np.where(self.src_pos[3, :] == self.rec_pos[3, ii])
                off = np.sqrt((self.rec_pos[0, ii] - self.src_pos[0, sid]) ** 2
                              + (self.rec_pos[1, ii] - self.src_pos[1, sid]) ** 2
                              + (self.rec_pos[2, ii] - self.src_pos[2, sid]) ** 2)
                if self.offmax >= off >= self.offmin:
                    valid.append(ii)
            self.rec_pos = self.rec_pos[:, valid]
            recels = [recels[el] for el in valid]
        self.rec_pos[4, :] = [x + 1 for x in recels]

    def ricker_wavelet(self, f0=None, NT=None, dt=None, tmin=None):
        NT = self.NT if NT is None else NT
        f0 = self.f0 if f0 is None else f0
        dt = self.dt if dt is None else dt
        tmin = -1.5 / f0 if tmin is None else tmin
        t = np.linspace(tmin, (NT - 1) * dt + tmin, num=int(NT))
        return ((1.0 - 2.0 * (np.pi ** 2) * (f0 ** 2) * (t ** 2))
                * np.exp(-(np.pi ** 2) * (f0 ** 2) * (t ** 2)))

    def surface_acquisition_2d(self, dg=2, ds=5, dsx=2, dsz=2, dgsz=0, src_type=100):
        self.src_pos_all = np.empty((5, 0))
        self.rec_pos_all = np.empty((8, 0))
        self.src_all = None
        nx = self.N[1]
        dlx = self.nab + dsx
        dlz = self.nab + dsz if self.freesurf == 0 else dsz
        gx = np.arange(dlx + dsx, nx - dlx, dg) * self.dh
        gz = gx * 0 + (dlz + dgsz) * self.dh
        for ii in range(dlx, nx - dlx, ds):
            idsrc = self.src_pos_all.shape[1]
            s = np.array([[ii * self.dh], [0], [dlz * self.dh], [idsrc], [src_type]], float)
            self.src_pos_all = np.append(self.src_pos_all, s, axis=1)
            gid = np.arange(0, len(gx)) + self.rec_pos_all.shape[1] + 1
            r = np.stack([gx, gz * 0, gz, gz * 0 + idsrc, gid, gx * 0, gx * 0, gx * 0], 0)
            self.rec_pos_all = np.append(self.rec_pos_all, r, axis=1)

    def crosshole_acquisition_2d(self, dg=2, ds=5, dsx=2, dsz=2):
        nz, nx = self.N
        dlx = self.nab + dsx
        dlz = self.nab + dsz if self.freesurf == 0 else dsz
        gz = np.arange(dlz, nz - dlz, dg) * self.dh
        gx = gz * 0 + (nx - dlx) * self.dh
        for ii in range(dlz, nz - dlz, ds):
            idsrc = self.src_pos_all.shape[1]
            s = np.array([[dlx * self.dh], [0], [ii * self.dh], [idsrc], [100]], float)
            self.src_pos_all = np.append(self.src_pos_all, s, axis=1)
            gid = np.arange(0, len(gx)) + self.rec_pos_all.shape[1] + 1
            r = np.stack([gx, gx * 0, gz, gz * 0 + idsrc, gid, gx * 0 + 2, gx * 0, gx * 0], 0)
            self.rec_pos_all = np.append(self.rec_pos_all, r, axis=1)

    # ---- the run (SeisCL.py:414-505) ---------------------------------------
    def set_forward(self, jobids, params, workdir=None, withgrad=True):
        self.prepare_data(jobids)
        self.gradout = 1 if withgrad else 0
        for p in self.params:
            if p not in params:
                raise SeisCLError(f"Parameter with {p} not defined\n")
        self.N = np.array(params[self.params[0]].shape, dtype=int)
        self._params = {k: np.asarray(v, np.float32) for k, v in params.items()}
        self._jobids = jobids
        if self.files:
            self.write_csts(workdir)
            self.write_model(params, workdir)

    def set_backward(self, residuals=None, workdir=None):
        """Adjoint sources given per output variable, (NT, ntraces) each, in
        the trace order of the prepared receivers (inputres mode)."""
        if residuals is not None:
            self._res_in = [np.asarray(r, np.float32) for r in residuals]
            self.inputres = 1
        else:
            self._res_in = None
            self.inputres = 0
        self.gradout = 1
        if self.files:
            if residuals is not None:
                h5mat.savemat(self._path(workdir, self.file_res),
                              {w + "res": r for w, r in zip(self.to_load_names, self._res_in)})
            self.write_csts(workdir)

    def write_data(self, data, workdir=None, filename=None):
        """Observed data for the residuals: {name: (NT, ntraces)} with the
        keys of to_load_names, in the trace order of the prepared receivers."""
        self._din = {k: np.asarray(v, np.float32) for k, v in data.items()}
        if self.files:
            data = dict(data)
            for k in ("src_pos", "rec_pos", "src"):
                if k not in data:
                    data[k] = getattr(self, k)
            h5mat.savemat(self._path(workdir, self.file_din if filename is None else filename),
                          data)

    # ---- files (SeisCL.py:682-763) ------------------------------------------
    def write_csts(self, workdir=None, filename=None):
        vals = {}
        for k in CSTS_KEYS:
            v = getattr(self, k)
            if v is None:
                raise SeisCLError(f"{k} is not set: call set_forward first")
            vals[k] = v
        # this engine's gradient corrections (absent from the reference's
        # files; SeisCL_MPI reads them, 0 reproduces the reference)
        vals["gradfix"] = int(self.gradfix)
        h5mat.savemat(self._path(workdir, self.file_csts if filename is None else filename),
                      vals)

    def write_model(self, params, workdir=None):
        for p in self.params:
            if p not in params:
                raise SeisCLError(f"Parameter with {p} not defined\n")
        h5mat.savemat(self._path(workdir, self.file_model),
                      {k: np.asarray(v).astype(np.float32) for k, v in params.items()})

    def callcmd(self, workdir=None):
        """The program and its arguments: base name and data file."""
        workdir = os.path.abspath(self.workdir if workdir is None else workdir)
        return [self.progname, os.path.join(workdir, self.file),
                os.path.join(workdir, self.file_din)]

    def _engine_csts(self):
        c = {k: getattr(self, k) for k in self._ENGINE_KEYS}
        c["gradfix"] = self.gradfix
        return c

    def execute(self, workdir=None):
        if self.files:
            cmd = self.callcmd(workdir)
            if not self.spawn:
                if h5mat.cli_main(cmd[1:]) != 0:
                    raise SeisCLError("SeisCL_MPI failed")
                return ""
            p = subprocess.run(cmd, capture_output=True)
            if p.stderr:
                raise SeisCLError(p.stderr.decode())
            if p.returncode:
                raise SeisCLError(f"SeisCL_MPI exited with {p.returncode}")
            return p.stdout.decode()
        if self._params is None:
            raise SeisCLError("call set_forward first")
        if self._ctx is None:
            self._ctx = Context(self.device)
        m = Model(self._ctx, self._params, **self._engine_csts())
        names = self.to_load_names
        if list(m.out_names) != list(names):
            m.close()
            raise SeisCLError(f"outputs {m.out_names} differ from to_load_names {names}")
        ntr = self.rec_pos.shape[1]
        out = {n: np.zeros((int(self.NT), ntr), np.float32) for n in names}
        res = {n: np.zeros((int(self.NT), ntr), np.float32) for n in names}
        rms = rmsnorm = 0.0
        grad = self.gradout == 1
        misfit = not self.inputres and (grad or self.rmsout or self.resout)
        nshot = len(set(self.src_pos[3, :].astype(int)))
        nframe = int(self.NT) // int(self.movout) if self.movout > 0 else 0
        # movie variables: output wavefields (the elastic pressure is derived
        # from the stresses and left out, as in the reference)
        mov = {n: np.zeros(tuple(int(v) for v in self.N) + (nframe, nshot), np.float32)
               for n in names if nframe and (n != "p" or self.ND == 22)}
        mov_shots = []
        if grad:
            if self.back_prop_type not in (1, 2):
                m.close()
                raise SeisCLError("back_prop_type must be 1 (backpropagation) or 2 (DFT)")
            if self.back_prop_type == 2:
                # gradient from the DFT of the wavefields at gradfreqs
                # (src/assign_modeling_case.c:549-570, src/calc_grad.c:377-968)
                if self.inputres or self.movout > 0:
                    m.close()
                    raise SeisCLError("inputres and movout need back_prop_type=1")
                m.grad_dft(np.asarray(self.gradfreqs, np.float32).ravel())
            m.grad_zero()
        if misfit and (self._din is None or any(n not in self._din for n in names)):
            m.close()
            raise SeisCLError(f"observed data must hold every output of {names}: "
                              "call write_data first")
        for sid in sorted(set(self.src_pos[3, :].astype(int))):
            scol = self.src_pos[3, :].astype(int) == sid
            rcol = np.where(self.rec_pos[3, :].astype(int) == sid)[0]
            shot = (self.src_pos[:, scol], self.src[:, scol], self.rec_pos[:, rcol])
            if grad and self.inputres and mov:
                # the given adjoint sources, scaled, and the gradient run's movie
                o = m.forward_save(*shot)
                r = m.res_scale([r[:, rcol] for r in self._res_in], shot[2])
                self._movie_frames(m.adjoint(r, scale=False, movout=int(self.movout),
                                             movie_vars=list(mov)), mov, len(mov_shots))
                mov_shots.append(sid)
            elif grad and self.inputres and sid in self._saved:
                # pass 2 of the inputres flow: the forward run saved by pass 1
                # (checkpoint_h2d, src/time_stepping.c:856-859); no seismograms
                m.saved_write(*shot, self._saved.pop(sid))
                r = m.res_scale([r[:, rcol] for r in self._res_in], shot[2])
                m.adjoint(r, scale=False)
                o = [np.zeros((int(self.NT), rcol.size), np.float32) for _ in names]
            elif grad and self.inputres:
                o, _ = m.gradient(*shot, res=[r[:, rcol] for r in self._res_in])
            elif self.inputres and self.movout <= 0:
                # pass 1 of the inputres flow: forward run with its saved state
                # kept for pass 2 (checkpoint_d2h, src/time_stepping.c:850-853)
                o = m.forward_save(*shot)
                self._saved[sid] = m.saved_read()
            elif grad:
                o = m.forward_save(*shot)
            elif self.movout > 0:
                o = self._forward_movie(m, shot, mov, len(mov_shots))
                mov_shots.append(sid)
            else:
                o = m.forward(*shot)
            if misfit:
                # var_res_raw with the misfit options, then res_scale: the
                # adjoint sources, which the reference also writes as <var>res
                din = [self._din[n][:, rcol] for n in names]
                r, (r1, r2) = m.residuals(o, din)
                rms += r1
                rmsnorm += r2
                r = m.res_scale(r, shot[2])
                if grad and mov:
                    # movie of the gradient run (src/time_stepping.c:926-929)
                    self._movie_frames(m.adjoint(r, scale=False, movout=int(self.movout),
                                                 movie_vars=list(mov)), mov, len(mov_shots))
                    mov_shots.append(sid)
                elif grad:
                    m.adjoint(r, scale=False)
                for n, a in zip(names, r):
                    res[n][:, rcol] = a
            for n, a in zip(names, o):
                out[n][:, rcol] = a
        self._results = {"out": out, "res": res, "rms": (rms, rmsnorm), "mov": mov}
        if grad:
            g = m.grad_read()
            self._results["grad"] = {n: g[n] for n in g}
            if self.Hout:
                h = m.hess_read()
                self._results["H"] = {n: h[n] for n in h}
        m.close()
        return ""

    @staticmethod
    def _movie_frames(frames, mov, k):
        """Frames of shot k from Model.adjoint's movie into the movie arrays."""
        for n, a in mov.items():
            a[..., k] = frames[n]

    def _forward_movie(self, m, shot, mov, k):
        """Forward shot k with a snapshot of each movie variable after every
        step t with (t + 1) % movout == 0 (src/time_stepping.c:272-350)."""
        m.bind_shot(*shot)
        H = int(self.FDORDER) // 2
        inner = tuple(slice(H, -H) for _ in range(len(self.N)))
        nt = int(self.NT)
        # the run stops at tmax, as seis_forward_shot and SeisCL_MPI's movie do
        # (times in seconds -> steps in float arithmetic, src/read_hdf5.c:523-526)
        tmax = int(np.float32(self.tmax) / np.float32(self.dt)) if self.tmax else nt
        t = 0
        for f in range(nt // int(self.movout)):
            t1 = (f + 1) * int(self.movout)
            if t1 > tmax:
                break
            m.run_steps(t, t1 - t)
            t = t1
            for n, a in mov.items():
                a[..., f, k] = m.read_field(n)[inner]
        if t < tmax:
            m.run_steps(t, tmax - t)
        return m.read_shot()

    # ---- outputs (SeisCL.py:507-656) ---------------------------------------
    def read_data(self, workdir=None, filename=None, residuals=False):
        if self.files:
            path = self._path(workdir, self.file_dout if filename is None else filename)
            suf = "res" if residuals else "out"
            try:
                have = set(h5mat.names(path))
            except Exception as e:
                raise SeisCLError("Could not read data") from e
            out = [np.transpose(h5mat.read(path, w + suf)) for w in self.to_load_names
                   if w + suf in have]
            if not out:
                raise SeisCLError("Could not read data: variables not found")
            return out
        key = "res" if residuals else "out"
        if key not in self._results:
            raise SeisCLError("Could not read data")
        return [self._results[key][n] for n in self.to_load_names]

    def _crop(self, o):
        if self.cropgrad:
            if self.freesurf == 0:
                o[:self.nab, :] = 0
            o[-self.nab:, :] = 0
            o[:, :self.nab] = 0
            o[:, -self.nab:] = 0
        return o

    def read_movie(self, workdir=None, filename=None):
        """Snapshots of the output wavefields every movout steps
        (SeisCL/SeisCL.py:543-570): per variable of to_load_names that is a
        wavefield, an array [NZ, (NY,) NX, NT // movout, nshots]."""
        if self.files:
            path = self._path(workdir, self.file_movout if filename is None else filename)
            try:
                have = set(h5mat.names(path))
            except Exception as e:
                raise SeisCLError("Could not read data") from e
            out = [np.transpose(h5mat.read(path, "mov" + w)) for w in self.to_load_names
                   if "mov" + w in have]
        else:
            mov = self._results.get("mov", {})
            out = [mov[w] for w in self.to_load_names if w in mov]
        if not out:
            raise SeisCLError("Could not read movie: variables not found")
        return out

    def read_grad(self, workdir=None, param_names=None, filename=None):
        if self.files:
            path = self._path(workdir, self.file_gout if filename is None else filename)
            param_names = self.params if param_names is None else param_names
            try:
                return [self._crop(np.transpose(h5mat.read(path, "grad" + n)).copy())
                        for n in param_names]
            except Exception as e:
                raise SeisCLError("Could not read grad") from e
        if "grad" not in self._results:
            raise SeisCLError("Could not read grad")
        param_names = self.params if param_names is None else param_names
        out = []
        for name in param_names:
            if name not in self._results["grad"]:
                raise SeisCLError(f"no gradient for {name}")
            out.append(self._crop(np.array(self._results["grad"][name])))
        return out

    def read_Hessian(self, workdir=None, param_names=None, filename=None):
        """The approximate Hessian written with the gradient when Hout=1
        (SeisCL/SeisCL.py:605-640): H<name> of the gradient file."""
        if param_names is None:
            param_names = self.params
        if self.files:
            path = self._path(workdir, self.file_gout if filename is None else filename)
            try:
                out = [np.transpose(h5mat.read(path, "H" + n)) for n in param_names]
            except Exception as e:
                raise SeisCLError("Could not read Hessian") from e
        else:
            if "H" not in self._results:
                raise SeisCLError("Could not read Hessian")
            out = [self._results["H"][n].copy() for n in param_names]
        return [self._crop(o) for o in out]

    def read_rms(self, workdir=None, filename=None):
        if self.files:
            path = self._path(workdir, self.file_rms if filename is None else filename)
            try:
                rms, norm = h5mat.read(path, "rms")[0], h5mat.read(path, "rms_norm")[0]
            except Exception as e:
                raise SeisCLError("Forward modeling failed, could not read rms\n") from e
            return rms / norm, norm
        if "rms" not in self._results or self._results["rms"][1] == 0:
            raise SeisCLError("Forward modeling failed, could not read rms\n")
        rms, norm = self._results["rms"]
        return rms / norm, norm

    def close(self):
        if self._ctx is not None:
            self._ctx.close()
            self._ctx = None
