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
    Function needs.  csts are SeisCL.py-style constants (as for Model)."""

    def __init__(self, ctx: Context, params: dict, **csts):
        self.model = Model(ctx, params, **csts)
        self.param_type = self.model.csts.param_type
        self.param_names = list(PARAM_NAMES.get(self.param_type, ("vp", "vs", "rho")))
        if self.model.csts.ND == 22:
            raise ValueError("gradients need the elastic case (ND 2 or 3)")
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
    @staticmethod
    def forward(ctx, prop: Propagator, shots, *params):
        p = {n: _np(t) for n, t in zip(prop.param_names, params)}
        prop.model.set_params(p)
        outs = []
        nout = len(prop.out_names)
        for sp, s, rp in shots:
            outs.extend(prop.model.forward_save(sp, s, rp))
        prop._token += 1
        prop._holder = prop._token
        ctx.prop = prop
        ctx.shots = shots
        ctx.token = prop._token
        ctx.nout = nout
        ctx.params = p
        ctx.device = params[0].device
        return tuple(torch.from_numpy(o).to(ctx.device) for o in outs)

    @staticmethod
    def backward(ctx, *grads):
        prop, shots, nout = ctx.prop, ctx.shots, ctx.nout
        m = prop.model
        # the model may hold other parameters or another shot's boundaries
        fresh = prop._holder == ctx.token
        if not fresh:
            m.set_params(ctx.params)
        m.grad_zero()
        for k in reversed(range(len(shots))):
            g = grads[k * nout:(k + 1) * nout]
            if all(x is None for x in g):
                continue
            sp, s, rp = shots[k]
            nt, nrec = m.csts.NT, np.asarray(rp).shape[1]
            res = [_np(x) if x is not None else np.zeros((nt, nrec), np.float32) for x in g]
            if not (fresh and k == len(shots) - 1):
                m.forward_save(sp, s, rp)
            m.adjoint(res)
            fresh = False
        prop._holder = None
        gr = m.grad_read()
        out = [torch.from_numpy(gr[n]).to(ctx.device) for n in prop.param_names]
        return (None, None) + tuple(out)


def seismic(prop: Propagator, shots, *params):
    """Seismograms of every shot, differentiable with respect to params.

    shots: list of (src_pos (5,ns), src (NT,ns), rec_pos (8,nr)) in
    SeisCL.py layout; params: tensors [NZ, (NY,) NX] in the order of
    prop.param_names.  Returns a flat tuple of (NT, nrec) tensors, shot by
    shot, each shot's outputs in prop.out_names order."""
    if len(params) != len(prop.param_names):
        raise ValueError(f"expected parameters {prop.param_names}")
    return _Seismic.apply(prop, list(shots), *params)
