"""GPU checks of the hand-written adjoints (nbh_force_vjp, nbh_adj_*) against
the FP64 torch reference (autograd of the same forward map, CPU) and against
central finite differences of GPU forward runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _banner(msg):
    print(f"[gpu-adjoint] {msg}", flush=True)


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


@pytest.fixture(scope="module")
def be(nb):
    with nb.Backend("fp64", 1.0 / 4096) as b:
        yield b


def unequal(nb, n, seed):
    s = nb.plummer(n, seed)
    s[3] *= np.random.default_rng(seed).lognormal(0.0, 0.1, n)
    return s


@pytest.mark.parametrize("eps2", [1.0 / 4096, 0.0])
def test_force_vjp_vs_torch_reference(nb, be, eps2):
    import torch
    import hermite_torch as ht
    n = 200
    s = unequal(nb, n, 41)
    cot = np.random.default_rng(3).normal(size=(6, n))
    be.configure(eps2=eps2)
    g = be.force_vjp(s, cot)
    x, v, m = (torch.from_numpy(s[0:3].copy()).requires_grad_(), torch.from_numpy(s[4:7].copy()).requires_grad_(),
               torch.from_numpy(s[3].copy()).requires_grad_())
    a, j = ht.force(x, v, m, eps2)
    ((a * torch.from_numpy(cot[0:3])).sum() + (j * torch.from_numpy(cot[3:6])).sum()).backward()
    ex, ev, em = rel(g[0:3], x.grad.numpy()), rel(g[4:7], v.grad.numpy()), rel(g[3], m.grad.numpy())
    _banner(f"force VJP eps2={eps2}: rel err x {ex:.1e} v {ev:.1e} m {em:.1e}")
    assert max(ex, ev, em) < 1e-12
    be.configure(eps2=1.0 / 4096)


def test_pair_force_gradcheck(nb, be):
    """torch.autograd.gradcheck of the ctypes autograd.Function (central
    differences of nbh_force vs nbh_force_vjp)."""
    import torch
    from nbody_amd.autograd import PairForce
    s = unequal(nb, 12, 5)
    be.configure(eps2=1e-2)
    x, v, m = (torch.from_numpy(s[0:3].copy()).requires_grad_(), torch.from_numpy(s[4:7].copy()).requires_grad_(),
               torch.from_numpy(s[3].copy()).requires_grad_())
    assert torch.autograd.gradcheck(lambda x, v, m: PairForce.apply(x, v, m, be), (x, v, m), eps=1e-6, atol=1e-7)
    be.configure(eps2=1.0 / 4096)


@pytest.mark.parametrize("n", [40, 128])
def test_trajectory_adjoint_vs_torch_reference(nb, be, n):
    """32 steps, checkpoints every 8: the final state of the forward run and
    the adjoint gradient of sum |x(T) - t|^2 + sum w.v(T) against the torch
    reference and its autograd.  n = 40 fits one source tile, so the force
    is evaluated in one split and the step kernel takes its planes as they
    are; n = 128 runs two splits."""
    import torch
    import hermite_torch as ht
    from nbody_amd.autograd import trajectory
    dt, nsteps, eps2 = 1.0 / 512, 32, 1.0 / 4096
    s = unequal(nb, n, 17)
    rng = np.random.default_rng(9)
    tgt = torch.from_numpy(s[0:3] + rng.normal(scale=0.01, size=(3, n)))
    wv = torch.from_numpy(rng.normal(size=(3, n)))
    be.configure(eps2=eps2)

    def loss(xT, vT):
        return ((xT - tgt) ** 2).sum() + (wv * vT).sum()

    leaves = [torch.from_numpy(s[0:3].copy()).requires_grad_(), torch.from_numpy(s[4:7].copy()).requires_grad_(),
              torch.from_numpy(s[3].copy()).requires_grad_()]
    xT, vT = trajectory(*leaves, be, dt, nsteps, 8)
    loss(xT, vT).backward()
    mine = [l.grad.numpy().copy() for l in leaves]
    ref_leaves = [l.detach().clone().requires_grad_() for l in leaves]
    rxT, rvT = ht.trajectory(*ref_leaves, dt, nsteps, eps2)
    loss(rxT, rvT).backward()
    ex, ev = rel(xT.detach().numpy(), rxT.detach().numpy()), rel(vT.detach().numpy(), rvT.detach().numpy())
    errs = [rel(a, b.grad.numpy()) for a, b in zip(mine, ref_leaves)]
    _banner(f"n={n} forward vs torch reference: rel err x {ex:.1e} v {ev:.1e}; adjoint: rel err x {errs[0]:.1e} "
            f"v {errs[1]:.1e} m {errs[2]:.1e}")
    assert max(ex, ev) < 1e-13
    assert max(errs) < 1e-10


def test_checkpoint_interval_does_not_change_gradient(nb, be):
    n, dt, nsteps = 256, 1.0 / 512, 24
    s = unequal(nb, n, 23)
    cot = np.zeros((7, n))
    cot[0:3] = np.random.default_rng(4).normal(size=(3, n))
    be.configure(eps2=1.0 / 4096)
    grads = []
    for every in (1, 5, 24, 100):
        with nb.Adjoint(be, n, every) as adj:
            adj.forward(s, dt, nsteps)
            grads.append(adj.backward(cot))
    for g in grads[1:]:
        assert np.array_equal(g, grads[0])


def differentiation_problem(nb, n=4096, seed=5, target_seed=6):
    """Plummer sphere (seed 5) with masses x lognormal(sigma=0.1) renormalised
    to total mass 1; targets are the positions after the same run with
    re-drawn masses (seed 6)."""
    s = nb.plummer(n, seed)
    m = s[3] * np.random.Generator(np.random.PCG64(seed)).lognormal(0.0, 0.1, n)
    s[3] = m / m.sum()
    mt = np.random.Generator(np.random.PCG64(target_seed)).lognormal(0.0, 0.1, n)
    st = s.copy()
    st[3] = mt / mt.sum()
    return s, st


def test_config5_adjoint_vs_central_differences(nb, be):
    """N=4096, eps=1/64, dt=1/512, 128 FP64 steps, checkpoints every 16;
    loss = sum |x_i(T) - x_i^target|^2; central differences (h = 1e-6) on 32
    random coordinates of (x0, v0, m)."""
    n, dt, nsteps, every, h = 4096, 1.0 / 512, 128, 16, 1e-6
    be.configure(eps2=(1.0 / 64) ** 2)
    s, st = differentiation_problem(nb, n)
    with nb.Adjoint(be, n, every) as adj:
        target = adj.forward(st, dt, nsteps)[0:3]

        def loss(state):
            return float(((adj.forward(state, dt, nsteps)[0:3] - target) ** 2).sum())

        fin = adj.forward(s, dt, nsteps)
        cot = np.zeros((7, n))
        cot[0:3] = 2.0 * (fin[0:3] - target)
        grad = adj.backward(cot)
        rng = np.random.default_rng(7)
        rows = rng.choice([0, 1, 2, 3, 4, 5, 6], 32)
        cols = rng.integers(0, n, 32)
        worst = 0.0
        for r, c in zip(rows, cols):
            sp, sm = s.copy(), s.copy()
            sp[r, c] += h
            sm[r, c] -= h
            fd = (loss(sp) - loss(sm)) / (2 * h)
            scale = max(abs(fd), 1e-3 * np.abs(grad[r]).max())
            worst = max(worst, abs(grad[r, c] - fd) / scale)
    _banner(f"config5 adjoint vs central differences (32 coords): worst rel err {worst:.2e}")
    assert worst < 1e-4


@pytest.mark.parametrize("n", [40, 1000])
def test_forward_steps_bitwise_those_of_the_resident_integrator(nb, n):
    """The adjoint's forward step (force at the prediction, then one kernel
    that adds the split sums, corrects and predicts the next step) against
    the resident integrator's k_predict, force, k_reduce_splits, k_correct,
    with the force launched in the same shape (LDS tiles of 64, one target
    per lane, the adjoint's j-split of about 8 workgroups per CU): the states
    after 24 steps are bitwise equal.  n = 40 is one split (the planes as the
    unsplit kernel writes them), n = 1000 sixteen."""
    dt, nsteps, eps2 = 1.0 / 512, 24, 1.0 / 4096
    cus = nb.device_info()["compute_units"]
    tiles = (n + 63) // 64
    splits = max(1, min(tiles, (8 * cus + tiles - 1) // tiles))
    s = unequal(nb, n, 31)
    with nb.Backend("fp64", eps2, block=64, jsplit=splits, ipl=1, sources="lds") as b:
        b.upload(s)
        b.set_dt(dt)
        b.step(nsteps)
        resident, _ = b.download()
        with nb.Adjoint(b, n, 8) as adj:
            fin = adj.forward(s, dt, nsteps)
    diff = np.abs(fin - resident).max()
    _banner(f"n={n}, {splits} splits: adjoint forward vs resident steps, max |difference| {diff:.1e}")
    assert np.array_equal(fin, resident)
