"""GPU parity of the device-resident Hermite integrator (predictor, force,
corrector, adaptive dt, energy, hipGraph replay) against the oracle and the
reference's golden trajectories, through the C ABI."""
import numpy as np
import pytest

from conftest import body_relerr, golden

pytestmark = pytest.mark.gpu

EPS2_REF = 1e-7


def _banner(msg):
    print(f"[gpu-integrator] {msg}", flush=True)


def oracle_trajectory(oracle, s, eps2, dt, nsteps):
    """Reference step order (src/nbody.cpp:224-289); yields (state, force)
    after the initial evaluation and after every step."""
    f0 = oracle.force(s, eps2)
    yield s, f0
    for _ in range(nsteps):
        p = oracle.predict(s, f0, dt)
        f1 = oracle.force(p, eps2)
        s, _, _ = oracle.snap_correct(p, f0, f1, dt)
        f0 = f1
        yield s, f0


def test_config1_energy_every_step(nb, oracle):
    """N=1024 Plummer seed 1, eps = 1/64, dt = 1/512, 256 FP64 steps.
    Gates: per-body a and jerk at the first two evaluations <= 1e-12; total
    energy at every step within 1e-12 relative of the oracle's.  The final
    per-body divergence is reported, not gated (the orbits are chaotic)."""
    s0 = nb.plummer(1024, 1)
    eps2, dt, nsteps = (1 / 64) ** 2, 1 / 512, 256
    traj = oracle_trajectory(oracle, s0, eps2, dt, nsteps)
    with nb.Backend("fp64", eps2) as b:
        b.upload(s0)
        b.set_dt(dt)
        worst_e = 0.0
        for k, (so, fo) in enumerate(traj):
            if k > 0:
                b.step(1)
            st, f = b.download()
            if k <= 1:
                assert body_relerr(f[:3], fo[:3]).max() <= 1e-12
                assert body_relerr(f[3:], fo[3:]).max() <= 1e-12
            kin, pot = b.energy()
            e_gpu = kin + pot
            e_or = oracle.kinetic(so) + oracle.potential(so, eps2, nthreads=16)
            worst_e = max(worst_e, abs(e_gpu - e_or) / abs(e_or))
        assert worst_e <= 1e-12, worst_e
        div = np.abs(st[0:3] - so[0:3]).max()
        _banner(f"config1: worst energy rel diff {worst_e:.2e}, final position divergence {div:.2e}")
        assert div < 1e-6


def test_fixed_dt_vs_reference_trajectory(nb):
    g = golden("plummer_n1024.npz")
    with nb.Backend("fp64", EPS2_REF) as b:
        b.upload(g["state"])
        b.set_dt(float(g["fixed_dt"]))
        b.step(8)
        st, f = b.download()
    assert np.abs(st - g["fixed_state"]).max() <= 1e-13
    assert body_relerr(f[:3], g["fixed_force"][:3]).max() <= 1e-10


def test_adaptive_dt_vs_reference_and_oracle(nb, oracle):
    """NBody::initialdt and updatedt (src/nbody.cpp:131-222).  The initial dt
    is checked against the reference trajectory; the updatedt sequence from
    the reference's own tiny first step is ill-conditioned (see
    test_oracle.py::test_adaptive_dt_ill_conditioned_at_reference_start), so
    it is gated from dt0 = 1/1024 against the oracle."""
    g = golden("plummer_n1024.npz")
    with nb.Backend("fp64", EPS2_REF) as b:
        b.upload(g["state"])
        assert b.initial_dt(1e-4) == pytest.approx(float(g["adapt_dts"][0]), rel=1e-12)
        b.set_dt(1 / 1024, adaptive=True, eta=1e-2)
        dts = [1 / 1024]
        for _ in range(8):
            b.step(1)
            dts.append(b.get_dt())
        st, _ = b.download()
    o = oracle.run(g["state"], EPS2_REF, 0.0, 8, adaptive=True, dt0=1 / 1024)
    assert np.allclose(dts[:8], o["dts"], rtol=1e-9, atol=0)
    assert dts[8] == pytest.approx(o["next_dt"], rel=1e-9)
    assert np.abs(st - o["state"]).max() <= 1e-12


@pytest.mark.parametrize("n", [2, 3, 8, 64])
def test_tiny_systems_vs_reference(nb, n):
    g = golden("tiny.npz")
    with nb.Backend("fp64", EPS2_REF) as b:
        b.upload(g[f"state{n}"])
        b.set_dt(float(g["dt"]))
        b.step(4)
        st, f = b.download()
    assert np.abs(st - g[f"final{n}"]).max() <= 1e-13
    assert body_relerr(f[:3], g[f"force{n}"][:3]).max() <= 1e-12


def test_graph_replay_matches_eager(nb):
    """Every replay of the captured step graph is checked against an eager
    twin, with eager work of the program (energy, downloads) in between."""
    s = nb.plummer(2000, 5)
    with nb.Backend("fp64", 1e-4) as g, nb.Backend("fp64", 1e-4) as e:
        for b in (g, e):
            b.upload(s)
            b.set_dt(1 / 1024)
        for r in range(5):
            g.step(32, graph=True)    # one replay of the 32-step graph
            e.step(32, graph=False)
            assert g.energy() == e.energy()
            sg, fg = g.download()
            se, fe = e.download()
            assert np.array_equal(sg, se) and np.array_equal(fg, fe), f"replay {r}"
        g.step(40, graph=True)        # one replay + 8 eager steps
        e.step(40)
        assert np.array_equal(g.download()[0], e.download()[0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_after_shape_change(nb, precision):
    """A graph captured after the launch shape changed (more j-splits, scalar
    sources) needs larger partial and record buffers than the eager steps
    before it: they are reserved before the capture, and every replay matches
    eager steps of the same shape bit for bit."""
    s = nb.plummer(3000, 12)
    with nb.Backend(precision, 1e-4) as g, nb.Backend(precision, 1e-4) as e:
        for b in (g, e):
            b.configure(block=64, ipl=1, jsplit=1, sources="lds")
            b.upload(s)
            b.set_dt(1 / 1024)
            b.step(4)
            b.configure(block=128, ipl=2, jsplit=24, sources="scalar")
        for r in range(3):
            g.step(8, graph=True)
            e.step(8, graph=False)
            assert np.array_equal(g.download()[0], e.download()[0]), f"replay {r}"


def test_adaptive_graph_matches_eager(nb):
    s = nb.plummer(1500, 6)
    res = []
    for graph in (False, True):
        with nb.Backend("fp64", 1e-4) as b:
            b.upload(s)
            b.set_dt(b.initial_dt(1e-4), adaptive=True)
            b.step(32, graph=graph)
            res.append((b.download()[0], b.get_dt()))
    assert np.array_equal(res[0][0], res[1][0])
    assert res[0][1] == res[1][1]


def test_energy_compat_vs_reference_fixture(nb):
    g = golden("ref_ic_n1024.npz")
    with nb.Backend("fp64", EPS2_REF) as b:
        kin, pot = b.energy_of(g["state"], compat=True)
    assert kin == pytest.approx(float(g["kinetic"]), rel=1e-14)
    assert pot == pytest.approx(float(g["potential"]), rel=1e-12)


def test_energy_physical_vs_oracle(nb, oracle):
    s = nb.plummer(3000, 13)
    s[3] *= np.random.default_rng(2).lognormal(0.0, 0.1, 3000)  # unequal masses
    for eps2 in (1e-4, 0.0):
        with nb.Backend("fp64", eps2) as b:
            kin, pot = b.energy_of(s)
        assert kin == pytest.approx(oracle.kinetic(s), rel=1e-13)
        assert pot == pytest.approx(oracle.potential(s, eps2, nthreads=16), rel=1e-12)


@pytest.mark.parametrize("n", [1, 2, 255, 256, 257, 513, 769, 1000, 1281, 2560])
def test_energy_ragged_sizes(nb, oracle, n):
    """Odd and even tile counts (the pair schedule takes each tile pair once)
    and partial last tiles, unequal masses, both potential definitions."""
    _banner(f"energy n={n}")
    s = nb.plummer(n, 40 + n)
    s[3] *= np.random.default_rng(n).lognormal(0.0, 0.3, n)
    with nb.Backend("fp64", 1e-4) as b:
        kin, pot = b.energy_of(s)
        kin_c, pot_c = b.energy_of(s, compat=True)
    assert kin == pytest.approx(oracle.kinetic(s), rel=1e-13)
    assert kin_c == kin
    if n == 1:
        assert pot == 0.0 and pot_c == 0.0
        return
    assert pot == pytest.approx(oracle.potential(s, 1e-4, nthreads=8), rel=1e-12)
    assert pot_c == pytest.approx(oracle.potential_compat(s), rel=1e-12)


def test_energy_million_bodies_properties(nb):
    """N = 2^20 (4096 tiles, an even ring): doubling every mass scales the
    kinetic energy by 2 and both potentials by 4 exactly (powers of two), and
    a permutation of the bodies moves the sums only by rounding.  A tile pair
    visited twice or missed would not survive the permutation."""
    n = 1 << 20
    s = nb.plummer(n, 3)
    perm = np.random.default_rng(5).permutation(n)
    s2 = s.copy()
    s2[3] *= 2.0
    with nb.Backend("fp64", 1e-6) as b:
        kin, pot = b.energy_of(s)
        kin2, pot2 = b.energy_of(s2)
        kinp, potp = b.energy_of(s[:, perm])
        _, potc = b.energy_of(s, compat=True)
        _, potc2 = b.energy_of(s2, compat=True)
    _banner(f"N=2^20 energy: kin {kin:.15e} pot {pot:.15e}; permuted pot rel diff {abs(potp - pot) / abs(pot):.1e}")
    assert kin2 == 2.0 * kin and pot2 == 4.0 * pot and potc2 == 4.0 * potc
    assert abs(kinp - kin) <= 1e-13 * kin
    assert abs(potp - pot) <= 1e-12 * abs(pot)


def test_mixed_energy_drift_config2(nb):
    """N=102400 Plummer seed 2, eps = 1/256, dt = 1/1024, 3 steps: the mixed
    form's relative energy drift stays within 10x of the FP64 drift (or below
    1e-10 when the FP64 drift itself is at round-off)."""
    s = nb.plummer(102400, 2)
    eps2, dt = (1 / 256) ** 2, 1 / 1024
    drift = {}
    for prec in ("fp64", "mixed"):
        with nb.Backend(prec, eps2) as b:
            b.upload(s)
            e0 = sum(b.energy())
            b.set_dt(dt)
            b.step(3)
            drift[prec] = abs(sum(b.energy()) - e0) / abs(e0)
    _banner(f"config2 energy drift fp64 {drift['fp64']:.3e} mixed {drift['mixed']:.3e}")
    assert drift["mixed"] <= max(10 * drift["fp64"], 1e-10)


def test_step_before_upload_is_state_error(nb):
    with nb.Backend("fp64", 1e-4) as b:
        with pytest.raises(nb.NBodyError) as e:
            b.step(1)
        assert e.value.code == nb.ERR_STATE
        b.upload(nb.plummer(64, 1))
        with pytest.raises(nb.NBodyError) as e:
            b.step(1)   # no dt yet
        assert e.value.code == nb.ERR_STATE


def test_hermite_fourth_order_circular_orbit(nb):
    """Two bodies of mass 1/2 on a circular orbit of unit separation (period
    2 pi, speed 1/2 each), unsoftened, one period in n fixed steps on the
    device: the return error |x(2 pi) - x(0)| of a fourth-order Hermite
    scheme falls as n^-4, so halving dt divides it by ~16.  A predictor or
    corrector term dropped or mis-weighted (src/nbody.cpp:152-196) makes the
    scheme lower order and the ratio 4 or 8; checked between 12 and 20 for
    n = 64 -> 128 -> 256, errors from ~1e-5 down."""
    s = np.zeros((7, 2))
    s[0] = [0.5, -0.5]
    s[3] = [0.5, 0.5]
    s[5] = [0.5, -0.5]
    errs = []
    for n in (64, 128, 256):
        with nb.Backend("fp64", 0.0) as b:
            b.upload(s)
            b.set_dt(2 * np.pi / n)
            b.step(n)
            st, _ = b.download()
        errs.append(np.abs(st[0:3] - s[0:3]).max())
    ratios = [errs[k] / errs[k + 1] for k in range(2)]
    _banner(f"circular orbit, one period: errors {['%.3e' % e for e in errs]}, ratios {['%.1f' % r for r in ratios]}")
    assert errs[0] < 1e-3
    assert all(12.0 <= r <= 20.0 for r in ratios), ratios
