"""Pins the test-side FP64 torch reference (oracle/hermite_torch.py) and the
step-adjoint algebra before they check the GPU adjoint.  CPU only."""
import numpy as np
import pytest

from conftest import body_relerr

def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a))


def test_torch_force_matches_oracle(oracle, nb):
    import hermite_torch as ht
    s = nb.plummer(300, 31)
    s[3] *= np.random.default_rng(1).lognormal(0, 0.1, 300)
    eps2 = 1.0 / 4096
    a, j = ht.force(_t(s[0:3]), _t(s[4:7]), _t(s[3]), eps2)
    ref = oracle.force(s, eps2)
    assert body_relerr(a.numpy(), ref[:3]).max() < 1e-13
    assert body_relerr(j.numpy(), ref[3:]).max() < 1e-13


def test_torch_trajectory_matches_oracle(oracle, nb):
    import hermite_torch as ht
    s = nb.plummer(128, 32)
    eps2, dt = 1.0 / 4096, 1.0 / 512
    x, v = ht.trajectory(_t(s[0:3]), _t(s[4:7]), _t(s[3]), dt, 16, eps2)
    o = oracle.run(s, eps2, dt, 16)
    assert np.abs(x.numpy() - o["state"][0:3]).max() < 1e-13
    assert np.abs(v.numpy() - o["state"][4:7]).max() < 1e-13


def test_folded_corrector_coefficients():
    """The adjoint uses x' = x_p - 0.15 dt^2 (a - a1) - 7/60 dt^3 j - 1/30 dt^3 j1,
    v' = v_p - 0.5 dt (a - a1) - 5/12 dt^2 j - 1/12 dt^2 j1; these must equal
    the reference's snap/crackle corrector (src/nbody.cpp:165-196)."""
    rng = np.random.default_rng(5)
    xp, vp, a, j, a1, j1 = rng.uniform(-1, 1, (6, 3, 50))
    dt = 0.0137
    a2 = 2.0 / dt ** 2 * (-3.0 * (a - a1) - (2.0 * j + j1) * dt)
    a3 = 6.0 / dt ** 3 * (2.0 * (a - a1) + (j + j1) * dt)
    x_ref = xp + a2 * dt ** 4 / 24 + a3 * dt ** 5 / 120
    v_ref = vp + a2 * dt ** 3 / 6 + a3 * dt ** 4 / 24
    x_fold = xp - 0.15 * dt ** 2 * (a - a1) - 7 / 60 * dt ** 3 * j - 1 / 30 * dt ** 3 * j1
    v_fold = vp - 0.5 * dt * (a - a1) - 5 / 12 * dt ** 2 * j - 1 / 12 * dt ** 2 * j1
    assert np.abs(x_fold - x_ref).max() < 1e-15
    assert np.abs(v_fold - v_ref).max() < 1e-14


def test_torch_force_gradient_vs_central_differences(nb):
    """The reference gradient itself: torch autograd of the dense force
    against central differences on a 10-body system."""
    import torch
    import hermite_torch as ht
    s = nb.plummer(10, 3)
    s[3] *= np.random.default_rng(2).lognormal(0, 0.1, 10)
    x, v, m = (_t(s[0:3]).clone().requires_grad_(), _t(s[4:7]).clone().requires_grad_(),
               _t(s[3]).clone().requires_grad_())
    assert torch.autograd.gradcheck(lambda x, v, m: ht.force(x, v, m, 1e-2), (x, v, m), eps=1e-6, atol=1e-7)
