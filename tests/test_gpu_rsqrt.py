"""The FP64 pair kernels' reciprocal square root, through the public API.

The reference computes 1/sqrt(s) with InvSqrt2D (src/utils.cpp:155-165),
which errs by up to 4.3e-16 relative.  The GPU kernels refine the hardware
v_rsq_f64 seed with one third-order step (about one ulp).  This test reaches
that device function through the ensemble: 100,000 independent two-body
systems, masses 1, at rest, separations |dx| log-uniform in [1e-6, 1e6]
along x with 26 significant bits (so s = dx^2 + 0 is exact in FP64), no
softening.  The acceleration of body 0 is then exactly sign(dx) / dx^2,
which numpy's long double gives to ~1e-19; the kernel's value is
r * r^2 * dx with r the refined rsqrt, so its relative error is at most
3 e_r + 1.5 ulp for an rsqrt error e_r.

Bar: max relative error <= 1e-15 (~9 ulp: e_r <= ~2.5 ulp).  A single
Newton step from the seed (e_r up to ~4e-15) would err by up to ~1.2e-14,
a factor ten above it; test_newton_step_would_fail restates that step on
the CPU and shows it fails the same bar."""
import numpy as np
import pytest

NSYS = 100_000


def separations(seed=17):
    rng = np.random.default_rng(seed)
    dx = np.exp(rng.uniform(np.log(1e-6), np.log(1e6), NSYS))
    m, e = np.frexp(dx)
    dx = np.ldexp(np.round(m * 2.0 ** 26) / 2.0 ** 26, e)  # 26 significant bits: dx^2 exact
    return dx * rng.choice([-1.0, 1.0], NSYS)


def exact_acc(dx):
    d = dx.astype(np.longdouble)
    return (np.sign(d) / (d * d)).astype(np.longdouble)


def relerr(got, dx):
    ref = exact_acc(dx)
    return np.abs((got.astype(np.longdouble) - ref) / ref).astype(np.float64)


def newton_step_acc(dx):
    """The kernel's acceleration with a single Newton step, y' = y + y e / 2,
    from a seed with the hardware's measured relative error (5.2e-8, worst
    case of either sign), in FP64 as pair_f64 evaluates it; e = 1 - s y^2 is
    formed in long double and rounded once, standing in for the fma."""
    s = dx * dx
    worst = np.zeros(dx.size)
    for d in (5.2e-8, -5.2e-8):
        y = (1.0 / np.sqrt(s)) * (1.0 + d)
        yy = y * y
        e = (1.0 - s.astype(np.longdouble) * yy).astype(np.float64)
        r = y + (y * e) * 0.5
        a = (1.0 * r * (r * r)) * dx
        worst = np.maximum(worst, relerr(a, dx))
    return worst


def test_separations_are_exact_squares():
    dx = separations()
    s = dx * dx
    assert np.all(s.astype(np.longdouble) == dx.astype(np.longdouble) ** 2)


def test_newton_step_would_fail():
    assert newton_step_acc(separations()).max() > 5e-15


@pytest.mark.gpu
def test_device_rsqrt_two_body_ensemble(nb):
    dx = separations()
    counts = np.full(NSYS, 2, dtype=np.int32)
    state = np.zeros((7, 2 * NSYS))
    state[0, 1::2] = dx
    state[3] = 1.0
    print(f"[gpu-rsqrt] {NSYS} two-body systems, |dx| in [{np.abs(dx).min():.1e}, {np.abs(dx).max():.1e}]",
          flush=True)
    with nb.Backend("fp64", pairing="directed") as b, nb.Ensemble(b, counts, 0.0) as e:
        e.upload(state)
        _, f = e.download()
    assert np.all(f[1:] == 0.0)  # ay, az and the jerk vanish: motion-free, along x
    err0 = relerr(f[0, 0::2], dx)
    err1 = relerr(-f[0, 1::2], dx)
    worst = max(err0.max(), err1.max())
    print(f"[gpu-rsqrt] acceleration rel err: max {worst:.2e}, mean {err0.mean():.2e}", flush=True)
    assert worst <= 1e-15
