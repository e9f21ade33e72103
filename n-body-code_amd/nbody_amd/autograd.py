"""torch.autograd.Function wrappers over the C ABI's hand-written adjoints.

CPU float64 tensors in and out; the work runs on the GPU through
libnbody_hip.so (nbh_force_vjp, nbh_adj_*).  They exist so gradients of the
backend can be checked with torch's machinery (gradcheck-style finite
differences, reference gradients); they are not a training framework.
"""
import numpy as np
import torch

from . import Adjoint


def _np(t):
    return t.detach().to(torch.float64).contiguous().numpy()


class PairForce(torch.autograd.Function):
    """(x (3,n), v (3,n), m (n,)) -> (a (3,n), jerk (3,n)) with nbh_force /
    nbh_force_vjp on `backend` (an FP64 nbody_amd.Backend)."""

    @staticmethod
    def forward(ctx, x, v, m, backend):
        state = np.concatenate([_np(x), _np(m)[None], _np(v)])
        f = backend.force(state)
        ctx.backend = backend
        ctx.state = state
        return torch.from_numpy(f[0:3].copy()), torch.from_numpy(f[3:6].copy())

    @staticmethod
    def backward(ctx, ga, gj):
        n = ctx.state.shape[1]
        ga = np.zeros((3, n)) if ga is None else _np(ga)
        gj = np.zeros((3, n)) if gj is None else _np(gj)
        g = ctx.backend.force_vjp(ctx.state, np.concatenate([ga, gj]))
        return (torch.from_numpy(g[0:3].copy()), torch.from_numpy(g[4:7].copy()),
                torch.from_numpy(g[3].copy()), None)


class HermiteTrajectory(torch.autograd.Function):
    """(x0, v0, m) -> (x(T), v(T)) after nsteps fixed-dt Hermite steps, with
    the checkpointed adjoint (nbh_adj_forward / nbh_adj_backward)."""

    @staticmethod
    def forward(ctx, x0, v0, m, adjoint, dt, nsteps):
        state0 = np.concatenate([_np(x0), _np(m)[None], _np(v0)])
        fin = adjoint.forward(state0, dt, nsteps)
        ctx.adjoint = adjoint
        return torch.from_numpy(fin[0:3].copy()), torch.from_numpy(fin[4:7].copy())

    @staticmethod
    def backward(ctx, gx, gv):
        n = ctx.adjoint.n
        cot = np.zeros((7, n))
        if gx is not None:
            cot[0:3] = _np(gx)
        if gv is not None:
            cot[4:7] = _np(gv)
        g = ctx.adjoint.backward(cot)
        return (torch.from_numpy(g[0:3].copy()), torch.from_numpy(g[4:7].copy()),
                torch.from_numpy(g[3].copy()), None, None, None)


def trajectory(x0, v0, m, backend, dt, nsteps, every):
    """Differentiable final (x, v); builds an nbh_adjoint of the right size."""
    adj = Adjoint(backend, x0.shape[1], every)
    return HermiteTrajectory.apply(x0, v0, m, adj, dt, nsteps)
