"""FP64 reference of the force and the Hermite trajectory in plain torch
(CPU), for reference gradients by torch autograd.

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the product's
hand-written adjoint kernels; never by the product.

The forward map restates src/kernel.cpp:1188-1205 (pair acceleration and
jerk, all j != i) and src/nbody.cpp:152-196 (predictor, snap/crackle,
corrector) with a fixed dt.  Dense (n, n) pair tensors: intended for
n up to a few hundred.  The forward values are pinned against the C oracle
in tests/test_adjoint_ref.py.
"""
import torch


def force(x, v, m, eps2):
    """x, v: (3, n); m: (n,) -> (a, jerk), each (3, n)."""
    n = x.shape[1]
    d = x[:, None, :] - x[:, :, None]          # d[c, i, j] = x_j - x_i
    u = v[:, None, :] - v[:, :, None]
    eye = torch.eye(n, dtype=torch.bool)
    s = (d * d).sum(0) + eps2
    s = torch.where(eye, torch.ones_like(s), s)
    w = torch.where(eye, torch.zeros_like(s), m[None, :].expand(n, n))
    r = s.rsqrt()
    q = r * r * r
    p = q * r * r
    du = (d * u).sum(0)
    a = (w * q * d).sum(2)
    jerk = (w * (q * u - 3.0 * du * p * d)).sum(2)
    return a, jerk


def step(x, v, a0, j0, m, dt, eps2):
    """One Hermite step in the reference's form; returns (x, v, a1, j1)."""
    xp = x + v * dt + 0.5 * a0 * dt * dt + j0 * dt ** 3 / 6.0
    vp = v + a0 * dt + 0.5 * j0 * dt * dt
    a1, j1 = force(xp, vp, m, eps2)
    a2 = 2.0 / dt ** 2 * (-3.0 * (a0 - a1) - (2.0 * j0 + j1) * dt)
    a3 = 6.0 / dt ** 3 * (2.0 * (a0 - a1) + (j0 + j1) * dt)
    xc = xp + a2 * dt ** 4 / 24.0 + a3 * dt ** 5 / 120.0
    vc = vp + a2 * dt ** 3 / 6.0 + a3 * dt ** 4 / 24.0
    return xc, vc, a1, j1


def trajectory(x0, v0, m, dt, nsteps, eps2):
    a, j = force(x0, v0, m, eps2)
    x, v = x0, v0
    for _ in range(nsteps):
        x, v, a, j = step(x, v, a, j, m, dt, eps2)
    return x, v
