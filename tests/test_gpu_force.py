"""GPU parity of the pair kernels against the oracle and the reference's
golden vectors, through the C ABI (libnbody_hip.so).

Bars: FP64 per-body |da|/|a| and |dj|/|j| <= 1e-12 (summation order differs
from the reference's j<i triangle, so not bitwise); mixed per-body median
relative acceleration error <= 1e-6.
"""
import numpy as np
import pytest

from conftest import body_relerr, golden

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
EPS2_REF = 1e-7


def _banner(msg):
    print(f"[gpu-force] {msg}", flush=True)


def check64(f, ref, tol=TOL64):
    ea = body_relerr(f[:3], ref[:3])
    ej = body_relerr(f[3:], ref[3:])
    assert ea.max() <= tol, f"acc max rel err {ea.max():.3e}"
    assert ej.max() <= tol, f"jerk max rel err {ej.max():.3e}"
    return ea.max(), ej.max()


@pytest.fixture(scope="module")
def be64(nb):
    with nb.Backend("fp64", EPS2_REF) as b:
        yield b


@pytest.fixture(scope="module")
def bemx(nb):
    with nb.Backend("mixed", EPS2_REF) as b:
        yield b


def test_fp64_vs_reference_fixtures(be64):
    for name in ("ref_ic_n1024.npz", "plummer_n1024.npz"):
        _banner(f"fixture {name}")
        g = golden(name)
        be64.configure(eps2=EPS2_REF)
        check64(be64.force(g["state"]), g["force64"])


def test_fp64_plummer_config1_vs_oracle(be64, oracle, nb):
    """N=1024 Plummer seed 1, eps = 1/64 (the softening is a parameter)."""
    s = nb.plummer(1024, 1)
    eps2 = (1.0 / 64) ** 2
    be64.configure(eps2=eps2)
    ea, ej = check64(be64.force(s), oracle.force(s, eps2))
    _banner(f"config1 fp64 acc {ea:.2e} jerk {ej:.2e}")


@pytest.mark.parametrize("n", [1, 2, 3, 63, 65, 257, 1000, 1023, 1025, 4097])
def test_fp64_ragged_sizes(be64, oracle, nb, n):
    _banner(f"ragged n={n}")
    s = nb.plummer(n, 1000 + n)
    be64.configure(eps2=1e-4, block=0, jsplit=0)
    f = be64.force(s)
    ref = oracle.force(s, 1e-4)
    if n == 1:
        assert np.all(f == 0.0)
    else:
        check64(f, ref)


@pytest.mark.parametrize("sources", ["lds", "scalar"])
@pytest.mark.parametrize("block", [64, 128, 256])
@pytest.mark.parametrize("jsplit", [1, 2, 7, 0])
@pytest.mark.parametrize("ipl", [1, 2])
def test_fp64_launch_shapes_agree(be64, oracle, nb, block, jsplit, ipl, sources):
    _banner(f"shape block={block} jsplit={jsplit} ipl={ipl} sources={sources}")
    s = nb.plummer(3000, 77)
    be64.configure(eps2=1e-3, block=block, jsplit=jsplit, ipl=ipl, sources=sources)
    try:
        assert be64.launch_shape(3000)["sources"] == sources
        check64(be64.force(s), oracle.force(s, 1e-3))
    finally:
        be64.configure(block=0, jsplit=0, ipl=0, sources="auto")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n,eps2", [(1, 1e-4), (65, 1e-4), (1000, 0.0), (3001, 1e-3), (20000, 1e-5)])
def test_source_paths_bitwise_identical(nb, precision, n, eps2):
    """LDS tiles and scalar loads of packed records feed each target the same
    sources in the same order: the forces agree bit for bit, for every shape."""
    s = nb.plummer(n, 90 + n)
    with nb.Backend(precision, eps2) as b:
        for block, ipl, jsplit in ((64, 1, 1), (128, 2, 3), (256, 1, 0)):
            b.configure(block=block, ipl=ipl, jsplit=jsplit, sources="lds")
            f_lds = b.force(s)
            b.configure(sources="scalar")
            f_sc = b.force(s)
            assert np.array_equal(f_lds, f_sc), (block, ipl, jsplit)


def test_default_source_path_by_size(nb):
    """Without tuning: LDS tiles for small n, scalar loads from 65536 (mixed)
    and 262144 (FP64) bodies; the symmetric kernel reports none."""
    with nb.Backend("mixed", 1e-4) as b:
        assert b.launch_shape(4096)["sources"] == "lds"
        assert b.launch_shape(65536)["sources"] == "scalar"
    with nb.Backend("fp64", 1e-4) as b:
        assert b.launch_shape(131072)["sources"] == "lds"
        assert b.launch_shape(262144)["sources"] == "scalar"
        b.configure(sources="lds")
        assert b.launch_shape(262144)["sources"] == "lds"
    with nb.Backend("mixed", 1e-4, pairing="symmetric") as b:
        assert b.launch_shape(4096)["sources"] == "none"


@pytest.mark.parametrize("ipl", [1, 2])
def test_mixed_targets_per_lane_identical(bemx, nb, ipl):
    """Each target sums its sources in the same order whatever the number of
    targets per lane, so IPL cannot change the mixed result."""
    s = nb.plummer(3001, 78)
    bemx.configure(eps2=1e-3, block=128, jsplit=3, ipl=1)
    ref = bemx.force(s)
    bemx.configure(ipl=ipl)
    try:
        assert np.array_equal(bemx.force(s), ref)
    finally:
        bemx.configure(block=0, jsplit=0, ipl=0)


def test_autotune_keeps_parity(be64, oracle, nb):
    s = nb.plummer(2500, 79)
    be64.configure(eps2=1e-3, block=0, jsplit=0, ipl=0)
    t = be64.autotune(2500)
    _banner(f"autotune n=2500: {t}")
    assert t["block"] in (64, 128, 256) and t["ipl"] in (1, 2)
    shape = be64.launch_shape(2500)
    assert (shape["block"], shape["ipl"]) == (t["block"], t["ipl"])
    check64(be64.force(s), oracle.force(s, 1e-3))


def test_fp64_deterministic(be64, nb):
    s = nb.plummer(5000, 3)
    be64.configure(eps2=1e-4, block=0, jsplit=0)
    assert be64.launch_shape(5000)["jsplit"] > 1
    assert np.array_equal(be64.force(s), be64.force(s))


def test_fp64_zero_softening(be64, oracle, nb):
    """eps2 = 0: the self pair is dropped, distinct bodies interact unsoftened."""
    s = nb.plummer(777, 8)
    be64.configure(eps2=0.0)
    check64(be64.force(s), oracle.force(s, 0.0))


def test_coincident_bodies_unsoftened_reported(be64, nb):
    s = nb.plummer(100, 9)
    s[0:3, 5] = s[0:3, 6]
    be64.configure(eps2=0.0)
    with pytest.raises(nb.NBodyError) as e:
        be64.force(s)
    assert e.value.code == nb.ERR_NONFINITE
    be64.configure(eps2=1e-4)
    be64.force(s)  # softened: finite


def test_mixed_vs_oracle(bemx, oracle, nb):
    s = nb.plummer(1024, 1)
    eps2 = (1.0 / 64) ** 2
    bemx.configure(eps2=eps2)
    f = bemx.force(s)
    ref = oracle.force(s, eps2)
    ea = body_relerr(f[:3], ref[:3])
    ej = body_relerr(f[3:], ref[3:])
    _banner(f"mixed acc median {np.median(ea):.2e} max {ea.max():.2e}; jerk median {np.median(ej):.2e}")
    assert np.median(ea) <= 1e-6
    assert ea.max() <= 1e-5
    assert np.median(ej) <= 1e-5


def check_mixed(f, ref):
    ea = body_relerr(f[:3], ref[:3])
    ej = body_relerr(f[3:], ref[3:])
    assert np.median(ea) <= 1e-6 and ea.max() <= 1e-5, f"acc median {np.median(ea):.2e} max {ea.max():.2e}"
    assert np.median(ej) <= 1e-5, f"jerk median {np.median(ej):.2e}"
    return np.median(ea), ea.max()


# the mixed directed kernel reads its sources from packed records at
# wave-uniform indices: ragged ends of fold runs, splits and targets
@pytest.mark.parametrize("sources", ["lds", "scalar"])
@pytest.mark.parametrize("n", [1, 2, 3, 63, 65, 257, 1000, 1023, 1025, 4097])
def test_mixed_ragged_sizes(bemx, oracle, nb, n, sources):
    _banner(f"mixed ragged n={n} sources={sources}")
    s = nb.plummer(n, 2000 + n)
    bemx.configure(eps2=1e-4, block=0, jsplit=0, ipl=0, sources=sources)
    f = bemx.force(s)
    bemx.configure(sources="auto")
    if n == 1:
        assert np.all(f == 0.0)
    else:
        check_mixed(f, oracle.force(s, 1e-4))


@pytest.mark.parametrize("sources", ["lds", "scalar"])
@pytest.mark.parametrize("block", [64, 128, 256])
@pytest.mark.parametrize("jsplit", [1, 2, 7, 0])
@pytest.mark.parametrize("ipl", [1, 2])
def test_mixed_launch_shapes_agree(bemx, oracle, nb, block, jsplit, ipl, sources):
    _banner(f"mixed shape block={block} jsplit={jsplit} ipl={ipl} sources={sources}")
    s = nb.plummer(3000, 77)
    bemx.configure(eps2=1e-3, block=block, jsplit=jsplit, ipl=ipl, sources=sources)
    try:
        check_mixed(bemx.force(s), oracle.force(s, 1e-3))
    finally:
        bemx.configure(block=0, jsplit=0, ipl=0, sources="auto")


@pytest.mark.parametrize("sources", ["lds", "scalar"])
def test_mixed_zero_softening(bemx, oracle, nb, sources):
    """eps2 = 0: the self pair is dropped in the mixed kernel too."""
    s = nb.plummer(777, 8)
    bemx.configure(eps2=0.0, sources=sources)
    f = bemx.force(s)
    bemx.configure(sources="auto")
    assert np.isfinite(f).all()
    check_mixed(f, oracle.force(s, 0.0))


def test_mixed_autotune_truncated_trials(bemx, oracle, nb):
    """n > 65536: the autotuner's first stage times a truncated source range
    (the packed records cover only those sources), the second all of them."""
    n = 70000
    s = nb.plummer(n, 81)
    bemx.configure(eps2=1e-4, block=0, jsplit=0, ipl=0)
    t = bemx.autotune(n)
    _banner(f"mixed autotune n={n}: {t}")
    f = bemx.force(s)
    idx = np.sort(np.random.default_rng(5).choice(n, 192, replace=False))
    check_mixed(f[:, idx], oracle.force_sample(s, 1e-4, idx, nthreads=16))


def test_mixed_vs_reference_f32_fixture(bemx):
    """The reference's _F32 build is FP32 arithmetic with double sums; the
    mixed kernel must sit within FP32 rounding of it."""
    g = golden("plummer_n1024.npz")
    bemx.configure(eps2=EPS2_REF)
    f = bemx.force(g["state"])
    e = body_relerr(f[:3], g["force32"][:3])
    assert np.median(e) <= 1e-6


@pytest.mark.parametrize("sources", ["lds", "scalar"])
def test_large_sample_vs_reference_fixture(be64, bemx, nb, sources):
    """N=102400 Plummer seed 2 (the 102,400-body configuration's sphere) at the
    reference's softening, every 100th body against the reference, with either
    source path."""
    g = golden("plummer_n102400_sample.npz")
    s = nb.plummer(int(g["n"]), int(g["seed"]))
    idx = g["idx"]
    try:
        be64.configure(eps2=EPS2_REF, sources=sources)
        f = be64.force(s)
        check64(f[:, idx], g["force64"])
        bemx.configure(eps2=EPS2_REF, sources=sources)
        fm = bemx.force(s)
    finally:
        be64.configure(sources="auto")
        bemx.configure(sources="auto")
    e = body_relerr(fm[:3, idx], g["force64"][:3])
    _banner(f"N=102400 mixed ({sources}) median {np.median(e):.2e}")
    assert np.median(e) <= 1e-6


def test_million_bodies_properties(be64, bemx, oracle, nb):
    """N = 1,048,576 Plummer seed 3, eps = 1e-3: a sample of bodies against the
    long-double directed sum, momentum conservation, mixed vs FP64."""
    n = 1 << 20
    s = nb.plummer(n, 3)
    be64.configure(eps2=1e-6)
    f = be64.force(s)
    idx = np.random.default_rng(0).choice(n, 24, replace=False)
    ref = oracle.force_sample(s, 1e-6, idx, nthreads=16)
    check64(f[:, idx], ref)
    m = s[3]
    for c in range(6):
        assert abs(np.sum(m * f[c])) <= 1e-12 * np.sum(np.abs(m * f[c]))
    bemx.configure(eps2=1e-6)
    fm = bemx.force(s)
    e = body_relerr(fm[:3], f[:3])
    _banner(f"N=1M mixed vs fp64 acc median {np.median(e):.2e} max {e.max():.2e}")
    assert np.median(e) <= 1e-6


def test_force_device_pointers(be64, oracle, nb):
    """nbh_force_device on caller-owned HBM buffers (allocated through the
    HIP runtime the library itself links)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    n = 2048
    s = np.ascontiguousarray(nb.plummer(n, 21))
    out = np.zeros((6, n))
    d_in, d_out = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(d_in), C.c_size_t(s.nbytes)) == 0
    assert hip.hipMalloc(C.byref(d_out), C.c_size_t(out.nbytes)) == 0
    try:
        assert hip.hipMemcpy(d_in, s.ctypes.data_as(C.c_void_p), C.c_size_t(s.nbytes), 1) == 0
        be64.configure(eps2=1e-4)
        col = lambda base, k: C.c_void_p(base.value + 8 * n * k)  # noqa: E731
        rc = nb.lib().nbh_force_device(be64._h, n, *[col(d_in, k) for k in range(7)],
                                       *[col(d_out, k) for k in range(6)], None)
        assert rc == nb.OK
        be64.timing()  # drains the context stream
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), d_out, C.c_size_t(out.nbytes), 2) == 0
    finally:
        hip.hipFree(d_in)
        hip.hipFree(d_out)
    check64(out, oracle.force(s, 1e-4))


def test_force_device_pointers_mixed(bemx, oracle, nb):
    """nbh_force_device in mixed precision with scalar-load sources: the
    context packs the caller's source arrays into its own record buffer,
    which grows with n."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    for n in (1500, 4100, 700):
        s = np.ascontiguousarray(nb.plummer(n, 30 + n))
        out = np.zeros((6, n))
        d_in, d_out = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(d_in), C.c_size_t(s.nbytes)) == 0
        assert hip.hipMalloc(C.byref(d_out), C.c_size_t(out.nbytes)) == 0
        try:
            assert hip.hipMemcpy(d_in, s.ctypes.data_as(C.c_void_p), C.c_size_t(s.nbytes), 1) == 0
            bemx.configure(eps2=1e-4, block=0, jsplit=0, ipl=0, sources="scalar")
            col = lambda base, k: C.c_void_p(base.value + 8 * n * k)  # noqa: E731
            rc = nb.lib().nbh_force_device(bemx._h, n, *[col(d_in, k) for k in range(7)],
                                           *[col(d_out, k) for k in range(6)], None)
            assert rc == nb.OK
            bemx.timing()  # drains the context stream
            assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), d_out, C.c_size_t(out.nbytes), 2) == 0
        finally:
            hip.hipFree(d_in)
            hip.hipFree(d_out)
        check_mixed(out, oracle.force(s, 1e-4))
        assert np.array_equal(out, bemx.force(s))  # same kernel through the host-array call
    bemx.configure(sources="auto")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_force_device_pointers_symmetric(oracle, nb, precision):
    """nbh_force_device with symmetric pairing: the atomic totals land in the
    caller's hipMalloc'd outputs (n = 2100: a partial last block)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    n = 2100
    s = np.ascontiguousarray(nb.plummer(n, 22))
    out = np.zeros((6, n))
    d_in, d_out = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(d_in), C.c_size_t(s.nbytes)) == 0
    assert hip.hipMalloc(C.byref(d_out), C.c_size_t(out.nbytes)) == 0
    try:
        assert hip.hipMemcpy(d_in, s.ctypes.data_as(C.c_void_p), C.c_size_t(s.nbytes), 1) == 0
        with nb.Backend(precision, 1e-4, pairing="symmetric") as b:
            assert b.launch_shape(n)["pairing"] == "symmetric"
            col = lambda base, k: C.c_void_p(base.value + 8 * n * k)  # noqa: E731
            rc = nb.lib().nbh_force_device(b._h, n, *[col(d_in, k) for k in range(7)],
                                           *[col(d_out, k) for k in range(6)], None)
            assert rc == nb.OK
            b.timing()  # drains the context stream
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), d_out, C.c_size_t(out.nbytes), 2) == 0
    finally:
        hip.hipFree(d_in)
        hip.hipFree(d_out)
    ref = oracle.force(s, 1e-4)
    if precision == "fp64":
        check64(out, ref)
    else:
        assert np.median(body_relerr(out[:3], ref[:3])) <= 1e-6


# eps2 > 0 but too small for the self pair's m eps2^-3/2 to stay finite in
# the pair arithmetic's type (FP32 below ~2e-26; FP64 below ~3e-206; 1e-40
# is an FP32 denormal, 1e-50 rounds to 0 in FP32): the self pair must be
# dropped explicitly there too, or inf * 0 turns every force into NaN.
@pytest.mark.parametrize("precision,eps2", [("mixed", 1e-30), ("mixed", 1e-40), ("mixed", 1e-50),
                                            ("fp64", 1e-200), ("fp64", 1e-300)])
@pytest.mark.parametrize("pairing,sources", [("directed", "lds"), ("directed", "scalar"), ("symmetric", "auto")])
def test_tiny_softening_drops_self_pair(nb, oracle, precision, eps2, pairing, sources):
    n = 2 * 768 + 300  # two symmetric blocks and a partial one
    _banner(f"tiny eps2 {precision} {eps2:g} {pairing} {sources}")
    s = nb.plummer(n, 91)
    with nb.Backend(precision, eps2, pairing=pairing, sources=sources) as b:
        assert b.launch_shape(n)["pairing"] == pairing
        f = b.force(s)  # raises ERR_NONFINITE if a NaN got through
    ref = oracle.force(s, eps2)
    if precision == "fp64":
        check64(f, ref)
    else:
        check_mixed(f, ref)


@pytest.mark.parametrize("precision,eps2", [("mixed", 1e-30), ("fp64", 1e-200)])
@pytest.mark.parametrize("pairing", ["directed", "symmetric"])
def test_tiny_softening_ensemble(nb, oracle, precision, eps2, pairing):
    counts = np.array([40, 600, 1100], dtype=np.int32)
    eps = np.array([1e-3, eps2, eps2])
    state = nb.ensemble_state(counts, 92)
    with nb.Backend(precision, pairing=pairing) as b, nb.Ensemble(b, counts, eps) as e:
        e.upload(state)
        _, f = e.download()
        assert np.isfinite(f).all()
        for k in range(len(counts)):
            fk, ref = e.system(f, k), oracle.force(e.system(state, k), eps[k])
            if precision == "fp64":
                check64(fk, ref)
            else:
                check_mixed(fk, ref)
