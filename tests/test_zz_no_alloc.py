"""nbh_force_device_noalloc: the force evaluation on caller-owned device
arrays with its working memory in the library's static device scratch
instead of context buffers.  Checked through the C ABI against
nbh_force_device on the same inputs, two calls in a row with the second
smaller, in both precisions and pairings; and the device's free memory is
unchanged across the calls."""
import ctypes as C

import numpy as np
import pytest

from conftest import body_relerr

pytestmark = pytest.mark.gpu


def _banner(msg):
    print(f"[gpu-noalloc] {msg}", flush=True)


class Dev:
    """A (rows, n) float64 array in HBM, allocated by the test through the
    HIP runtime the library links."""

    def __init__(self, hip, rows, n):
        self.hip, self.rows, self.n = hip, rows, n
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(8 * rows * n)) == 0

    def col(self, k):
        return C.c_void_p(self.p.value + 8 * self.n * k)

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (self.rows, self.n)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def get(self):
        a = np.empty((self.rows, self.n))
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(a.nbytes), 2) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


@pytest.fixture(scope="module")
def hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return h


def free_bytes(hip):
    f, t = C.c_size_t(), C.c_size_t()
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
    return f.value


def evaluate(nb, hip, b, state, entry):
    """One evaluation through `entry` on device copies of state; returns the
    (6, n) forces and the device's free bytes just before and after it."""
    n = state.shape[1]
    din, dout = Dev(hip, 7, n), Dev(hip, 6, n)
    try:
        din.put(state)
        dout.put(np.full((6, n), np.nan))
        before = free_bytes(hip)
        rc = getattr(nb.lib(), entry)(b._h, n, *[din.col(k) for k in range(7)],
                                      *[dout.col(k) for k in range(6)], None)
        assert rc == nb.OK, nb.lib().nbh_last_error().decode()
        after = free_bytes(hip)
        return dout.get(), before, after
    finally:
        din.free()
        dout.free()


CASES = [
    # precision, pairing, config, sizes (second smaller), bitwise equal to nbh_force_device
    ("fp64", "directed", {"jsplit": 4, "sources": "lds"}, (5000, 1200), True),
    ("fp64", "directed", {"jsplit": 3, "sources": "scalar"}, (4100, 777), True),
    ("mixed", "directed", {"jsplit": 4, "sources": "lds"}, (5000, 1200), True),
    ("mixed", "directed", {"jsplit": 5, "sources": "scalar"}, (4100, 777), True),
    ("fp64", "symmetric", {}, (4000, 1700), False),
    ("mixed", "symmetric", {}, (4000, 1700), False),
]


@pytest.mark.parametrize("precision,pairing,cfg,sizes,bitwise", CASES)
def test_noalloc_matches_force_device(nb, hip, precision, pairing, cfg, sizes, bitwise):
    with nb.Backend(precision, 1e-4, pairing=pairing, **cfg) as b:
        for call, n in enumerate(sizes):
            s = nb.plummer(n, 500 + n)
            _banner(f"{precision} {pairing} {cfg} call {call} n={n}")
            ref, _, _ = evaluate(nb, hip, b, s, "nbh_force_device")
            got, before, after = evaluate(nb, hip, b, s, "nbh_force_device_noalloc")
            assert np.isfinite(got).all()
            if bitwise:
                assert np.array_equal(got, ref)
            else:  # FP64 atomic totals: the same sums in another order
                tol = 1e-12 if precision == "fp64" else 1e-9
                assert body_relerr(got[:3], ref[:3]).max() <= tol
                assert body_relerr(got[3:], ref[3:]).max() <= tol
            if call > 0:  # kernels loaded by now: the call itself takes no device memory
                assert after == before, (before, after)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_narrows_large_automatic_shapes(nb, hip, oracle, precision):
    """n = 65536 with the automatic shape: scalar-load sources and ~64
    j-splits (~200 MB of partials) do not fit the 64 MiB scratch, so the call
    runs fewer splits (then LDS sources); the result equals
    nbh_force_device's to FP64 rounding of the split sums (FP64) or to the
    mixed kernels' FP32 run sums (mixed), and both meet the bar against the
    oracle on a sample.  A second, smaller call follows."""
    with nb.Backend(precision, 1e-4) as b:
        for call, n in enumerate((65536, 30000)):
            s = nb.plummer(n, 600 + call)
            _banner(f"{precision} automatic shape n={n}: {b.launch_shape(n)}")
            ref, _, _ = evaluate(nb, hip, b, s, "nbh_force_device")
            got, before, after = evaluate(nb, hip, b, s, "nbh_force_device_noalloc")
            idx = np.sort(np.random.default_rng(n).choice(n, 128, replace=False))
            o = oracle.force_sample(s, 1e-4, idx, nthreads=8)
            ea, ej = body_relerr(got[:3], ref[:3]), body_relerr(got[3:], ref[3:])
            if precision == "fp64":
                assert ea.max() <= 1e-12 and ej.max() <= 1e-12
                assert body_relerr(got[:3, idx], o[:3]).max() <= 1e-12
            else:
                assert np.median(ea) <= 1e-6 and ea.max() <= 1e-5
                assert np.median(body_relerr(got[:3, idx], o[:3])) <= 1e-6
            if call > 0:
                assert after == before, (before, after)


def test_noalloc_keeps_a_forced_source_path_and_narrows_the_split(nb, hip):
    """LDS sources forced, j-split automatic, n = 65536: the call keeps the
    LDS path and runs fewer splits; FP64 sums equal nbh_force_device's to
    rounding of the split order."""
    n = 65536
    with nb.Backend("fp64", 1e-4, sources="lds") as b:
        s = nb.plummer(n, 8)
        ref, _, _ = evaluate(nb, hip, b, s, "nbh_force_device")
        got, _, _ = evaluate(nb, hip, b, s, "nbh_force_device_noalloc")
    assert body_relerr(got[:3], ref[:3]).max() <= 1e-12
    assert body_relerr(got[3:], ref[3:]).max() <= 1e-12


def test_noalloc_rejects_forced_shape_beyond_scratch(nb, hip):
    """A forced j-split whose partial sums exceed the 64 MiB scratch is
    rejected with an argument error before anything runs, and the outputs
    are untouched."""
    n = 65536
    with nb.Backend("fp64", 1e-4, jsplit=64, sources="lds") as b:
        din, dout = Dev(hip, 7, n), Dev(hip, 6, n)
        try:
            din.put(nb.plummer(n, 7))
            dout.put(np.full((6, n), 3.0))
            rc = nb.lib().nbh_force_device_noalloc(b._h, n, *[din.col(k) for k in range(7)],
                                                   *[dout.col(k) for k in range(6)], None)
            assert rc == nb.ERR_ARG
            assert b"scratch" in nb.lib().nbh_last_error()
            assert np.all(dout.get() == 3.0)
        finally:
            din.free()
            dout.free()
