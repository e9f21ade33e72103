"""The pair formula of the reference (src/kernel.cpp:1188-1205) against its
closed form for two bodies, through the public API.

20,000 independent two-body systems in one ensemble, no softening, unit
masses: body 0 at rest at the origin, body 1 at (dx, 0, 0) moving either
across the separation, (0, w, 0), or along it, (w, 0, 0).  Separations are
log-uniform in [1e-3, 1e3] and speeds in [1e-2, 1e2], with 26 significant
bits so that dx^2 and dx * w are exact in FP64.  For body 0:
  a = sign(dx) / dx^2 along x;
  jerk = (u - 3 (d.u) d / |d|^2) / |d|^3 = w / |dx|^3 along y (across),
         -2 w / |dx|^3 along x (along): the two terms of the jerk;
body 1 gets the negatives (the third law).  numpy's long double gives these
to ~1e-19.

Bars, per body and component that is not zero: FP64 1e-15 (acceleration),
2e-15 (jerk across, one more rounding than the acceleration), 5e-15 (jerk
along, where u - 3 (d.u) d/|d|^2 = -2u cancels two thirds of its second
term); the components that must vanish are exactly zero.  Mixed precision:
5e-7 for the acceleration and the jerk across (FP32 roundings of d and u,
~1.5 ulp of v_rsq_f32 tripled in r^3, a few products); 2e-6 for the jerk
along, where the FP32 error of the second term (-3 (d.u) r^2 d, some 6e-7
from r^2, d.u and their products) is carried by a result a third of its
size times two, i.e. grows by 1.5, on top of r^3's; the zero components
again exact."""
import numpy as np
import pytest

NSYS = 20_000


def bits26(x):
    m, e = np.frexp(x)
    return np.ldexp(np.round(m * 2.0 ** 26) / 2.0 ** 26, e)


def systems(seed=29):
    rng = np.random.default_rng(seed)
    dx = bits26(np.exp(rng.uniform(np.log(1e-3), np.log(1e3), NSYS))) * rng.choice([-1.0, 1.0], NSYS)
    w = bits26(np.exp(rng.uniform(np.log(1e-2), np.log(1e2), NSYS))) * rng.choice([-1.0, 1.0], NSYS)
    along = rng.random(NSYS) < 0.5
    return dx, w, along


def test_inputs_are_exact_in_fp64():
    dx, w, _ = systems()
    L = np.longdouble
    assert np.all((dx * dx).astype(L) == dx.astype(L) ** 2)
    assert np.all((dx * w).astype(L) == dx.astype(L) * w.astype(L))


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", (1e-15, 2e-15, 5e-15)), ("mixed", (5e-7, 5e-7, 2e-6))])
def test_two_body_closed_form(nb, precision, tol):
    dx, w, along = systems()
    counts = np.full(NSYS, 2, dtype=np.int32)
    state = np.zeros((7, 2 * NSYS))
    state[0, 1::2] = dx
    state[3] = 1.0
    state[4, 1::2] = np.where(along, w, 0.0)
    state[5, 1::2] = np.where(along, 0.0, w)
    print(f"[gpu-closed-form] {precision}: {NSYS} two-body systems, {int(along.sum())} moving along the "
          f"separation", flush=True)
    with nb.Backend(precision, pairing="directed") as b, nb.Ensemble(b, counts, 0.0) as e:
        e.upload(state)
        _, f = e.download()
    L = np.longdouble
    d, u = dx.astype(L), w.astype(L)
    r3 = np.abs(d) ** 3
    acc = np.sign(d) / (d * d)
    jerk_across = u / r3
    jerk_along = -2 * u / r3
    # body 0 (even rows), body 1 (odd rows): a and jerk of opposite sign
    for body, sgn in ((0, 1.0), (1, -1.0)):
        fa = f[:, body::2]
        ea = np.abs((fa[0].astype(L) - sgn * acc) / acc).astype(float)
        assert ea.max() <= tol[0], f"acc body {body}: {ea.max():.2e}"
        assert np.all(fa[1] == 0.0) and np.all(fa[2] == 0.0)
        ja = np.abs((fa[3][along].astype(L) - sgn * jerk_along[along]) / jerk_along[along]).astype(float)
        jc = np.abs((fa[4][~along].astype(L) - sgn * jerk_across[~along]) / jerk_across[~along]).astype(float)
        print(f"[gpu-closed-form] {precision} body {body}: acc {ea.max():.2e}, jerk across {jc.max():.2e}, "
              f"along {ja.max():.2e}", flush=True)
        assert jc.max() <= tol[1], f"jerk across body {body}: {jc.max():.2e}"
        assert ja.max() <= tol[2], f"jerk along body {body}: {ja.max():.2e}"
        assert np.all(fa[4][along] == 0.0) and np.all(fa[3][~along] == 0.0) and np.all(fa[5] == 0.0)
