"""GPU parity away from the easy cases, through the C ABI: a system far from
the origin, massless bodies, rotated inputs, a bulk velocity and exact
power-of-two rescaling, for the directed and the symmetric kernels in both
precisions.

The far-from-origin case pins that the mixed kernels form x_j - x_i in FP64
and only then round it to FP32, as the reference's _F32 build does
(src/kernel.cpp:1127-1133).  With the system moved ~1e5 from the origin, a
difference formed from FP32-rounded positions is off by up to ulp32(1e5) =
7.8e-3; test_fp32_differences_would_fail_far_from_origin restates that
mistake on the CPU and shows it fails the mixed bar by orders of magnitude,
so the GPU test can tell the two apart.

Bars: FP64 per-body |da|/|a| and |dj|/|j| <= 1e-12 against the oracle;
mixed median |da|/|a| <= 1e-6 (max <= 1e-5) and median |dj|/|j| <= 1e-5.
"""
import numpy as np
import pytest

from conftest import body_relerr

N = 4096  # at least two symmetric blocks (512 bodies) in both precisions
EPS2 = 1e-4
OFFSET = np.array([1.0e5, -2.5e4, 7.0e4])

KERNELS = [("directed", "lds"), ("directed", "scalar"), ("symmetric", "auto")]


def _banner(msg):
    print(f"[gpu-invariance] {msg}", flush=True)


def check(precision, f, ref):
    ea = body_relerr(f[:3], ref[:3])
    ej = body_relerr(f[3:], ref[3:])
    if precision == "fp64":
        assert ea.max() <= 1e-12, f"acc max rel err {ea.max():.3e}"
        assert ej.max() <= 1e-12, f"jerk max rel err {ej.max():.3e}"
    else:
        assert np.median(ea) <= 1e-6 and ea.max() <= 1e-5, f"acc median {np.median(ea):.2e} max {ea.max():.2e}"
        assert np.median(ej) <= 1e-5, f"jerk median {np.median(ej):.2e}"
    return np.median(ea), ea.max()


def force(nb, precision, pairing, sources, s, eps2=EPS2):
    with nb.Backend(precision, eps2, pairing=pairing, sources=sources) as b:
        shape = b.launch_shape(s.shape[1])
        assert shape["pairing"] == pairing
        if pairing == "directed":
            assert shape["sources"] == sources
        return b.force(s)


@pytest.fixture(scope="module")
def far(nb, oracle):
    s = nb.plummer(N, 31)
    s[0:3] += OFFSET[:, None]
    return s, oracle.force(s, EPS2, nthreads=8)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("pairing,sources", KERNELS)
def test_far_from_origin(nb, far, precision, pairing, sources):
    s, ref = far
    _banner(f"far from origin {precision} {pairing} {sources}")
    med, mx = check(precision, force(nb, precision, pairing, sources, s), ref)
    _banner(f"  acc median {med:.2e} max {mx:.2e}")


def test_fp32_differences_would_fail_far_from_origin(far):
    """The mistake the far-from-origin test catches, on the CPU: positions
    rounded to FP32 before the subtraction (everything after in FP64) miss
    the mixed bar by orders of magnitude, while differences formed in FP64
    and then rounded meet it."""
    s, ref = far
    idx = np.arange(0, N, 61)

    def acc_of(d):  # d: (3, targets, sources) position differences
        r2 = (d ** 2).sum(0) + EPS2
        mr3 = s[3][None, :] / (r2 * np.sqrt(r2))
        mr3[np.arange(idx.size), idx] = 0.0
        return (mr3[None] * d).sum(2)

    x32 = s[0:3].astype(np.float32)
    wrong = acc_of((x32[:, None, :] - x32[:, idx, None]).astype(np.float64))
    right = acc_of((s[0:3][:, None, :] - s[0:3][:, idx, None]).astype(np.float32).astype(np.float64))
    e_wrong = np.median(body_relerr(wrong, ref[:3, idx]))
    e_right = np.median(body_relerr(right, ref[:3, idx]))
    assert e_wrong > 1e-3, f"FP32-formed differences only err {e_wrong:.2e}"
    assert e_right <= 1e-7, f"FP64-formed, rounded differences err {e_right:.2e}"


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("pairing,sources", KERNELS)
def test_massless_bodies(nb, oracle, precision, pairing, sources):
    """Every seventh body massless (test particles), and n with a partial
    symmetric block: massless sources add nothing, massless targets are
    still accelerated."""
    n = N + 100
    s = nb.plummer(n, 32)
    s[3, ::7] = 0.0
    ref = oracle.force(s, EPS2, nthreads=8)
    _banner(f"massless {precision} {pairing} {sources}")
    check(precision, force(nb, precision, pairing, sources, s), ref)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("pairing,sources", KERNELS)
def test_rotation_equivariance(nb, precision, pairing, sources):
    """F(R s) = R F(s) for an orthogonal R applied to positions and
    velocities (no oracle involved).  FP64: to 1e-10, the rounding of the
    rotated inputs amplified by close pairs; mixed: the mixed bar."""
    s = nb.plummer(N, 33)
    q, r = np.linalg.qr(np.random.default_rng(33).standard_normal((3, 3)))
    R = q * np.sign(np.diag(r))
    t = s.copy()
    t[0:3] = R @ s[0:3]
    t[4:7] = R @ s[4:7]
    _banner(f"rotation {precision} {pairing} {sources}")
    f = force(nb, precision, pairing, sources, s)
    g = force(nb, precision, pairing, sources, t)
    rf = np.vstack([R @ f[0:3], R @ f[3:6]])
    ea = body_relerr(g[:3], rf[:3])
    ej = body_relerr(g[3:], rf[3:])
    if precision == "fp64":
        assert ea.max() <= 1e-10 and ej.max() <= 1e-10, f"acc {ea.max():.2e} jerk {ej.max():.2e}"
    else:
        assert np.median(ea) <= 1e-6 and np.median(ej) <= 1e-5, \
            f"acc median {np.median(ea):.2e} jerk median {np.median(ej):.2e}"


# A bulk velocity ~2e3 times the spheres' internal speeds (velocity
# dispersion ~0.4 in Henon units).
BOOST = np.array([1.0e3, -4.0e2, 7.0e2])


def jerk_of(s, idx, eps2, u_of):
    """CPU jerk of targets idx of s with FP64-formed position differences
    rounded to FP32 (the mixed kernels' d) and velocity differences from
    u_of(idx) -> (3, targets, sources); everything else in FP64."""
    d = (s[0:3][:, None, :] - s[0:3][:, idx, None]).astype(np.float32).astype(np.float64)
    u = u_of(idx)
    r2 = (d ** 2).sum(0) + eps2
    r = 1.0 / np.sqrt(r2)
    mr3 = s[3][None, :] * r ** 3
    mr3[np.arange(idx.size), idx] = 0.0
    al = -3.0 * (d * u).sum(0) * r * r
    return (mr3[None] * (u + al[None] * d)).sum(2)


@pytest.fixture(scope="module")
def boosted(nb, oracle):
    s = nb.plummer(N, 34)
    t = s.copy()
    t[4:7] += BOOST[:, None]
    return s, t, oracle.force(t, EPS2, nthreads=8)


def test_fp32_velocities_would_fail_under_a_boost(boosted):
    """The mistake the bulk-velocity test catches, on the CPU: velocities
    rounded to FP32 before the subtraction (as the reference's _F32 build
    does not: it forms v_j - v_i in FP64, src/kernel.cpp:1131-1133) lose the
    jerk once the system moves, while velocities taken relative to the
    centre-of-mass velocity in FP64 and then rounded keep it."""
    s, t, ref = boosted
    idx = np.arange(0, N, 37)
    v32 = t[4:7].astype(np.float32)
    vc = (t[3] * t[4:7]).sum(1) / t[3].sum()
    w32 = (t[4:7] - vc[:, None]).astype(np.float32)
    wrong = jerk_of(t, idx, EPS2, lambda i: (v32[:, None, :] - v32[:, i, None]).astype(np.float64))
    right = jerk_of(t, idx, EPS2, lambda i: (w32[:, None, :] - w32[:, i, None]).astype(np.float64))
    e_wrong = np.median(body_relerr(wrong, ref[3:, idx]))
    e_right = np.median(body_relerr(right, ref[3:, idx]))
    assert e_wrong > 1e-5, f"FP32-rounded velocities only err {e_wrong:.2e}"
    assert e_right <= 1e-6, f"centre-of-mass-relative velocities err {e_right:.2e}"


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("pairing,sources", KERNELS)
def test_bulk_velocity(nb, boosted, precision, pairing, sources):
    """Galilean invariance: the same sphere at rest and moving at BOOST.
    Forces depend on velocity differences only, so both must meet the bar
    against the oracle of the moving sphere; in mixed precision the jerk must
    also keep its at-rest accuracy (median <= 1e-6, against ~2e-5 with
    FP32-rounded velocities, test_fp32_velocities_would_fail_under_a_boost)
    and agree with the at-rest evaluation."""
    s, t, ref = boosted
    _banner(f"bulk velocity {precision} {pairing} {sources}")
    f = force(nb, precision, pairing, sources, t)
    check(precision, f, ref)
    g = force(nb, precision, pairing, sources, s)
    ej = body_relerr(f[3:], ref[3:])
    dj = body_relerr(f[3:], g[3:])
    da = body_relerr(f[:3], g[:3])
    _banner(f"  jerk vs oracle median {np.median(ej):.2e}; moving vs at rest: acc max {da.max():.2e}, "
            f"jerk median {np.median(dj):.2e} max {dj.max():.2e}")
    if precision == "fp64":
        assert da.max() <= 1e-12 and dj.max() <= 1e-9
    else:
        assert np.median(ej) <= 1e-6, f"jerk median {np.median(ej):.2e}"
        assert da.max() <= 1e-6 and np.median(dj) <= 1e-7, (da.max(), np.median(dj))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("pairing,sources", KERNELS)
def test_power_of_two_scale_covariance(nb, precision, pairing, sources):
    """Positions scaled by 2^-10, velocities by 2^5 and eps2 by 2^-20: every
    scaling is exact in floating point, and the pair formula is homogeneous,
    so accelerations must scale by exactly 2^20 and jerks by 2^35 (x^-2 and
    x^-3 v).  The directed kernels, whose sums have a fixed order, agree bit
    for bit (the hardware reciprocal square roots scale exactly by powers of
    4); the symmetric kernels' FP64 atomic sums may add in another order, so
    they are checked to 1e-13 relative per body."""
    s = nb.plummer(N, 37)
    t = s.copy()
    t[0:3] *= 2.0 ** -10
    t[4:7] *= 2.0 ** 5
    f = force(nb, precision, pairing, sources, s)
    g = force(nb, precision, pairing, sources, t, eps2=EPS2 * 2.0 ** -20)
    g[0:3] *= 2.0 ** -20
    g[3:6] *= 2.0 ** -35
    ea = body_relerr(g[:3], f[:3])
    ej = body_relerr(g[3:], f[3:])
    _banner(f"{precision} {pairing}/{sources}: scale covariance max rel acc {ea.max():.1e}, jerk {ej.max():.1e}, "
            f"bitwise {np.array_equal(g, f)}")
    assert ea.max() <= 1e-13 and ej.max() <= 1e-13
    if pairing == "directed":
        assert np.array_equal(g, f)
