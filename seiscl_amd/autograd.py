"""torch.autograd.Function over the HIP forward and adjoint runs.

``seismic(prop, shots, vp, vs, rho)`` returns the seismograms of every shot
as torch tensors; their backward pass is the reference's gradient by
backpropagation (SeisCL with gradout=1, back_prop_type=1), with the incoming
output gradients as adjoint sources (the inputres mode, SeisCL/SeisCL.py
set_backward with given residuals).  For L = 1/2 sum (out - d)^2 the output
gradient is the residual out - d, so ``L.backward()`` yields the misfit
gradient with respect to the model parameters.

Checkpointing is boundary-saving recomputation: the forward pass keeps the
boundary shell of the last shot resident on the GPU; the backward pass runs
the adjoint of that shot directly and recomputes the forward run (with
boundary saving) of every other shot before its adjoint.
"""
from __future__ import annotations

import numpy as np
import torch

from .engine import PARAM_NAMES, Context, Model


class Propagator:
    """A model resident on one GPU plus the bookkeeping the autograd
    Function needs.  csts are SeisCL.py-style constants (as for Model).

    ``param_names`` are the differentiable parameters: the three of
    param_type 0-2 (vp, vs, rho / M, mu, rho / Ip, Is, rho) and, for a
    viscoelastic model (L > 0), taup and taus.  seismic() also accepts only
    the first three for a viscoelastic model; taup and taus then keep the
    values the Propagator was created with.

    The gradient is the reference's correlation formula with the
    corrections of csts["gradfix"] (default 223).  gradfix=255 adds bit 32,
    which chains the vs and rho terms through the staggered averages of the
    parameter transform: the backward pass is then the derivative of the
    discrete misfit (finite differences agree to 1e-4)."""

    def __init__(self, ctx: Context, params: dict, **csts):
        self.model = Model(ctx, params, **csts)
        c = self.model.csts
        self.param_type = c.param_type
        if c.ND == 22:
            raise ValueError("gradients need the elastic case (ND 2 or 3)")
        if self.param_type not in PARAM_NAMES:
            raise ValueError("gradients need param_type 0, 1 or 2")
        base = list(PARAM_NAMES[self.param_type])
        self.visco = c.L > 0
        self.param_names = base + (["taup", "taus"] if self.visco else [])
        self._fixed_tau = ({n: np.asarray(params[n], np.float32) for n in ("taup", "taus")}
                           if self.visco else {})
        self._token = 0
        self._holder = None       # token of the forward pass whose last shot is saved

    @property
    def out_names(self):
        return self.model.out_names

    def close(self):
        self.model.close()


def _np(t: torch.Tensor) -> np.ndarray:
    return np.ascontiguousarray(t.detach().to("cpu", torch.float32).numpy())


class _Seismic(torch.autograd.Function):
    """Forward: every shot with boundary saving; backward: the adjoint runs
    (recomputing the forward run of every shot but the last).  With the
    parameters on the GPU, parameters, seismograms, adjoint sources and
    gradients stay in device memory (the C-ABI's *_device entry points);
    with CPU tensors the same runs go through host arrays."""

    @staticmethod
    def forward(ctx, prop: Propagator, shots, *params):
        names = prop.param_names[:len(params)]
        dev = params[0].is_cuda
        m = prop.model
        if dev:
            p = dict(zip(names, params))
            for n in prop.param_names[len(params):]:
                p[n] = torch.from_numpy(prop._fixed_tau[n]).to(params[0].device)
            p = {n: t.detach() for n, t in p.items()}
            m.set_params_device(p)
        else:
            p = {n: _np(t) for n, t in zip(names, params)}
            for n in prop.param_names[len(params):]:
                p[n] = prop._fixed_tau[n]
            m.set_params(p)
        outs = []
        nout = len(prop.out_names)
        for sp, s, rp in shots:
            if dev:
                o = m.forward_save_device(sp, s, rp)
                outs.extend(o[q].t().contiguous() for q in range(nout))
            else:
                outs.extend(torch.from_numpy(o) for o in m.forward_save(sp, s, rp))
        prop._token += 1
        prop._holder = prop._token
        ctx.prop = prop
        ctx.shots = shots
        ctx.token = prop._token
        ctx.nout = nout
        ctx.params = p
        ctx.names = names
        ctx.dev = dev
        ctx.device = params[0].device
        return tuple(o.to(ctx.device) for o in outs)

    @staticmethod
    def backward(ctx, *grads):
        prop, shots, nout, dev = ctx.prop, ctx.shots, ctx.nout, ctx.dev
        m = prop.model
        # the model may hold other parameters or another shot's boundaries
        fresh = prop._holder == ctx.token
        if not fresh:
            (m.set_params_device if dev else m.set_params)(ctx.params)
        m.grad_zero()
        for k in reversed(range(len(shots))):
            g = grads[k * nout:(k + 1) * nout]
            if all(x is None for x in g):
                continue
            sp, s, rp = shots[k]
            nt, nrec = m.csts.NT, np.asarray(rp).shape[1]
            if not (fresh and k == len(shots) - 1):
                if dev:
                    m.forward_save_device(sp, s, rp)
                else:
                    m.forward_save(sp, s, rp)
            if dev:
                res = torch.stack([x.t() if x is not None else
                                   torch.zeros((nrec, nt), device=ctx.device) for x in g])
                m.adjoint_device(res)
            else:
                res = [_np(x) if x is not None else np.zeros((nt, nrec), np.float32) for x in g]
                m.adjoint(res)
            fresh = False
        prop._holder = None
        if dev:
            gr = m.grad_read_device()
            out = [gr[n] for n in ctx.names]
        else:
            gr = m.grad_read()
            out = [torch.from_numpy(gr[n]).to(ctx.device) for n in ctx.names]
        return (None, None) + tuple(out)


def seismic(prop: Propagator, shots, *params):
    """Seismograms of every shot, differentiable with respect to params.

    shots: list of (src_pos (5,ns), src (NT,ns), rec_pos (8,nr)) in
    SeisCL.py layout; params: tensors [NZ, (NY,) NX] in the order of
    prop.param_names (for a viscoelastic model, either all five or only the
    first three), all on the CPU or all on the model's GPU.  Returns a flat
    tuple of (NT, nrec) tensors on the parameters' device, shot by shot,
    each shot's outputs in prop.out_names order."""
    n = len(params)
    if n != len(prop.param_names) and not (prop.visco and n == 3):
        raise ValueError(f"expected parameters {prop.param_names}")
    if len({t.device for t in params}) != 1:
        raise ValueError("all parameters must be on one device")
    return _Seismic.apply(prop, list(shots), *params)
