"""Symmetric (Newton's third law) pairing: every unordered pair evaluated once
and applied to both bodies, as the reference's j < i loops
(src/kernel.cpp:1126,1187).  Checked against the FP64 oracle (mixed to its
tolerance, FP64 to the drop-in's 1e-12), against the directed kernels, and
through properties the third law fixes."""
import numpy as np
import pytest

from conftest import body_relerr

pytestmark = pytest.mark.gpu

# per-body relative error bars against the FP64 oracle:
# (acc median, acc max, jerk median, jerk max)
TOL = {"mixed": (1e-6, 2e-5, 1e-5, 2e-4), "fp64": (1e-12, 1e-12, 1e-12, 1e-12)}


def _banner(msg):
    print(f"[gpu-sym] {msg}", flush=True)


def force(nb, state, eps2, pairing, precision="mixed"):
    with nb.Backend(precision, eps2, pairing=pairing) as b:
        shape = b.launch_shape(state.shape[1])
        return b.force(state), shape


def block_of(nb, precision):
    """Bodies per block of the symmetric kernel (threads x targets per lane)."""
    with nb.Backend(precision, 1e-4, pairing="symmetric") as b:
        shape = b.launch_shape(1 << 16)
    assert shape["pairing"] == "symmetric" and shape["block"] * shape["ipl"] % 64 == 0
    return shape["block"] * shape["ipl"]


# 2, 3, 4, 5 full blocks of B (even and odd ring); then a partial last block of
# 1, 76, 200, 300 bodies after 2 full blocks and 452, 441 after 3: a source
# block only, of which its first 1, 2, 4, 5, 8, 7 tiles of 64 are visited, the
# last one padded inside
SIZES = [(2, 0), (3, 0), (4, 0), (5, 0), (2, 1), (2, 76), (2, 200), (2, 300), (3, 452), (3, 441)]


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("blocks,rest", SIZES)
def test_symmetric_matches_oracle(nb, oracle, blocks, rest, precision):
    n = blocks * block_of(nb, precision) + rest
    _banner(f"oracle {precision} n={n}")
    s = nb.plummer(n, 11 + n)
    eps2 = (1 / 64) ** 2
    f, shape = force(nb, s, eps2, "symmetric", precision)
    assert shape["pairing"] == "symmetric"
    ref = oracle.force(s, eps2)
    ea = body_relerr(f[:3], ref[:3])
    ej = body_relerr(f[3:], ref[3:])
    _banner(f"{precision} n={n}: acc median {np.median(ea):.2e} max {ea.max():.2e}; "
            f"jerk median {np.median(ej):.2e} max {ej.max():.2e}")
    am, ax, jm, jx = TOL[precision]
    assert np.median(ea) <= am and ea.max() <= ax
    assert np.median(ej) <= jm and ej.max() <= jx


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_symmetric_close_to_directed(nb, precision):
    n = 8192
    _banner(f"directed vs symmetric {precision} n={n}")
    s = nb.plummer(n, 5)
    eps2 = 1e-6
    fs, _ = force(nb, s, eps2, "symmetric", precision)
    fd, shape = force(nb, s, eps2, "directed", precision)
    assert shape["pairing"] == "directed"
    ea = body_relerr(fs[:3], fd[:3])
    ej = body_relerr(fs[3:], fd[3:])
    _banner(f"{precision}: acc median {np.median(ea):.2e} max {ea.max():.2e}; "
            f"jerk median {np.median(ej):.2e} max {ej.max():.2e}")
    am, ax, jm, jx = TOL[precision]
    assert np.median(ea) <= am and np.median(ej) <= jm
    if precision == "fp64":
        assert ea.max() <= 1e-12 and ej.max() <= 1e-12


# 399 full blocks and a partial one of 100 bodies: workgroups take several
# entries of a block's source list, which ends with the partial block
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_partial_block_inside_workgroup_ranges(nb, oracle, precision):
    B = block_of(nb, precision)
    n = 399 * B + 100
    _banner(f"partial block, multi-pair workgroups {precision} n={n}")
    s = nb.plummer(n, 23)
    eps2 = (1 / 256) ** 2
    fs, shape = force(nb, s, eps2, "symmetric", precision)
    assert shape["pairing"] == "symmetric"
    fd, _ = force(nb, s, eps2, "directed", precision)
    ea = body_relerr(fs[:3], fd[:3])
    ej = body_relerr(fs[3:], fd[3:])
    _banner(f"{precision} vs directed: acc median {np.median(ea):.2e} max {ea.max():.2e}; "
            f"jerk median {np.median(ej):.2e} max {ej.max():.2e}")
    # a sample from every block kind: the partial one, its neighbours, the first
    idx = np.unique(np.concatenate([np.arange(n - 100, n), np.arange(n - 100 - B, n - 100, 8),
                                    np.arange(0, B, 8), np.random.default_rng(9).choice(n, 64)]))
    ref = oracle.force_sample(s, eps2, idx, nthreads=16)
    eo = body_relerr(fs[:3, idx], ref[:3])
    eoj = body_relerr(fs[3:, idx], ref[3:])
    am, ax, jm, jx = TOL[precision]
    assert np.median(eo) <= am and eo.max() <= ax
    assert np.median(eoj) <= jm and eoj.max() <= jx
    assert np.median(ea) <= am and np.median(ej) <= jm
    if precision == "fp64":
        assert ea.max() <= 1e-12 and ej.max() <= 1e-12


@pytest.mark.parametrize("precision,rel", [("mixed", 1e-6), ("fp64", 1e-13)])
def test_third_law_momentum(nb, precision, rel):
    """sum_i m_i a_i = 0 and sum_i m_i jerk_i = 0: each pair adds equal and
    opposite terms, so the totals vanish to the rounding of the sums."""
    n = 4000
    s = nb.plummer(n, 7)
    s[3] *= np.exp(0.3 * np.random.default_rng(3).standard_normal(n))
    f, _ = force(nb, s, 1e-5, "symmetric", precision)
    m = s[3]
    for c in range(6):
        tot = np.sum(m * f[c])
        scale = np.sum(np.abs(m * f[c]))
        assert abs(tot) <= rel * scale, (c, tot, scale)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("rest", [0, 52])
def test_eps2_zero_diagonal_pass(nb, oracle, precision, rest):
    """eps2 = 0: the self pair is dropped in the block-diagonal pass, and the
    pad bodies of a partial block stay away from every body."""
    n = 3 * block_of(nb, precision) + rest
    s = nb.plummer(n, 13)
    f, _ = force(nb, s, 0.0, "symmetric", precision)
    assert np.isfinite(f).all()
    ref = oracle.force(s, 0.0)
    assert np.median(body_relerr(f[:3], ref[:3])) <= TOL[precision][0]


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_unsupported_n_runs_directed(nb, precision):
    B = block_of(nb, precision)
    for n in (2 * B - 1, B, 1000, 7):
        s = nb.plummer(n, 2)
        fs, shape = force(nb, s, 1e-4, "symmetric", precision)
        assert shape["pairing"] == "directed"
        fd, _ = force(nb, s, 1e-4, "directed", precision)
        assert np.array_equal(fs, fd)
    # the symmetric kernels address the outputs with 32-bit byte offsets, so
    # they take n <= 2^29; above that the directed kernel is selected (the
    # launch-shape query allocates nothing)
    with nb.Backend(precision, 1e-4, pairing="symmetric") as b:
        assert b.launch_shape(1 << 29)["pairing"] == "symmetric"
        assert b.launch_shape((1 << 29) + 1)["pairing"] == "directed"


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_repeat_agrees_to_fp64_rounding(nb, precision):
    n = 8192
    s = nb.plummer(n, 17)
    with nb.Backend(precision, 1e-6, pairing="symmetric") as b:
        f1 = b.force(s)
        f2 = b.force(s)
    d = np.abs(f1 - f2).max(axis=1) / np.abs(f1).max(axis=1)
    _banner(f"{precision} repeat max rel diff per component {d.max():.2e}")
    assert d.max() <= 1e-13


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_large_sample_vs_oracle(nb, oracle, precision):
    n = 102400
    _banner(f"{precision} n={n} sample")
    s = nb.plummer(n, 2)
    eps2 = (1 / 256) ** 2
    f, shape = force(nb, s, eps2, "symmetric", precision)
    assert shape["pairing"] == "symmetric"
    idx = np.sort(np.random.default_rng(4).choice(n, 256, replace=False))
    ref = oracle.force_sample(s, eps2, idx, nthreads=16)
    ea = body_relerr(f[:3, idx], ref[:3])
    ej = body_relerr(f[3:, idx], ref[3:])
    _banner(f"{precision}: acc median {np.median(ea):.2e} max {ea.max():.2e}; jerk max {ej.max():.2e}")
    if precision == "mixed":
        assert np.median(ea) <= 1e-6 and ea.max() <= 1e-5
    else:
        assert ea.max() <= 1e-12 and ej.max() <= 1e-12


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("n", [4096, 4100])
def test_resident_steps_and_graph(nb, precision, n):
    """The integrator on the symmetric kernel: eager and graph replay agree to
    the atomics' rounding, and the energy drift matches the directed kernel's."""
    s = nb.plummer(n, 21)
    eps2 = (1 / 64) ** 2
    res = {}
    for pairing, graph in (("directed", True), ("symmetric", False), ("symmetric", True)):
        with nb.Backend(precision, eps2, pairing=pairing) as b:
            b.upload(s)
            e0 = sum(b.energy())
            b.set_dt(1 / 512)
            b.step(64, graph=graph)
            e1 = sum(b.energy())
            res[(pairing, graph)] = (b.download()[0], abs(e1 - e0) / abs(e0))
    xs_e, ds_e = res[("symmetric", False)]
    xs_g, ds_g = res[("symmetric", True)]
    xd, dd = res[("directed", True)]
    dx = np.abs(xs_e[:3] - xd[:3]).max()
    _banner(f"{precision}: drift directed {dd:.2e} symmetric {ds_e:.2e} / graph {ds_g:.2e}; "
            f"max |x_sym - x_dir| {dx:.2e}")
    assert np.abs(xs_e - xs_g).max() <= 1e-10
    assert dx <= (1e-6 if precision == "mixed" else 1e-11)
    assert ds_e <= max(10 * dd, 1e-10)


def test_million_bodies_symmetric_vs_directed(nb, oracle):
    """N = 2^20 (1365 full blocks of 768 and a partial one of 256 bodies): FP64
    symmetric per body to 1e-12 against the oracle's sum on a sample (the
    partial block, its neighbour and random bodies) and to 2e-12 against the
    FP64 directed kernel (two sums in different orders, each within 1e-12),
    mixed symmetric against FP64 to the mixed bar, and the third law's
    momentum sum."""
    n = 1 << 20
    s = nb.plummer(n, 3)
    eps2 = 1e-6
    fd, _ = force(nb, s, eps2, "directed", "fp64")
    fs, shape = force(nb, s, eps2, "symmetric", "fp64")
    assert shape["pairing"] == "symmetric"
    ea = body_relerr(fs[:3], fd[:3])
    ej = body_relerr(fs[3:], fd[3:])
    idx = np.unique(np.concatenate([np.arange(n - 8, n), np.arange(n - 300, n - 292),
                                    np.random.default_rng(1).choice(n, 16, replace=False)]))
    ref = oracle.force_sample(s, eps2, idx, nthreads=16)
    eo = body_relerr(fs[:3, idx], ref[:3])
    eoj = body_relerr(fs[3:, idx], ref[3:])
    fm, _ = force(nb, s, eps2, "symmetric", "mixed")
    em = body_relerr(fm[:3], fd[:3])
    _banner(f"N=2^20 fp64 sym vs oracle sample acc {eo.max():.2e} jerk {eoj.max():.2e}; vs dir acc max "
            f"{ea.max():.2e} jerk max {ej.max():.2e}; mixed sym acc median {np.median(em):.2e} max {em.max():.2e}")
    assert eo.max() <= 1e-12 and eoj.max() <= 1e-12
    assert ea.max() <= 2e-12 and ej.max() <= 2e-12
    assert np.median(em) <= 1e-6 and em.max() <= 1e-5
    m = s[3]
    for c in range(3):
        assert abs(np.sum(m * fs[c])) <= 1e-13 * np.sum(np.abs(m * fs[c]))
