"""The reference's host utility API (include/utils.h:29,55-63) as the repo's
host side offers it (n-body-code_amd/csrc/nbody.hpp, nbody_utils.cpp):
code written against utils.h links and gets the same values.

CPU: a small program linked like nbody_app computes InvSqrt2D and InvSqrt2
on 20,000 inputs, checked bit for bit against the reference's own functions
(oracle/_ref, compiled from src/utils.cpp) and, for InvSqrt2D, against the
oracle; timer() reads wall-clock seconds; estimate_err is |En-E0|/|E0|.
GPU: nbody_app_debug runs the reference's _DEBUG check test() after every
force evaluation; its host brute-force forces must agree with the GPU's."""
import os
import re
import subprocess
import time

import numpy as np
import pytest

from conftest import ROOT

PKG = os.path.join(ROOT, "n-body-code_amd")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hu") / "host_utils_check")
    src = [os.path.join(ROOT, "tests", "host_utils_check.cpp")] + \
          [os.path.join(PKG, "csrc", f) for f in ("nbody_utils.cpp", "kernel_hip.cpp")]
    lib = os.path.join(PKG, "lib")
    if not os.path.exists(os.path.join(lib, "libnbody_hip.so")):
        pytest.fail("libnbody_hip.so is not built")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                        "-I" + os.path.join(PKG, "csrc"), *src, "-L" + lib, "-lnbody_hip",
                        "-Wl,-rpath," + lib, "-o", out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return out


def run_checker(checker, tokens):
    r = subprocess.run([checker], input=" ".join(tokens), capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout.split("\n")


def test_invsqrt_bitwise_and_timer(checker, oracle, reference):
    ref64, _ = reference
    rng = np.random.default_rng(12)
    xd = np.exp(rng.uniform(np.log(1e-300), np.log(1e300), 10000))
    xf = np.exp(rng.uniform(np.log(1e-37), np.log(1e37), 10000)).astype(np.float32)
    toks = [f"{v:016x}" for v in xd.view(np.uint64)] + [f"{v:08x}" for v in xf.view(np.uint32)]
    t_before = time.time()
    lines = run_checker(checker, toks)
    got_d = np.array([int(l, 16) for l in lines[:10000]], dtype=np.uint64).view(np.float64)
    got_f = np.array([int(l, 16) for l in lines[10000:20000]], dtype=np.uint32).view(np.float32)
    want_d = np.array([ref64.invsqrt2d(x) for x in xd])
    want_f = np.array([ref64.invsqrt2(x) for x in xf], dtype=np.float32)
    assert np.array_equal(got_d.view(np.uint64), want_d.view(np.uint64))
    assert np.array_equal(got_f.view(np.uint32), want_f.view(np.uint32))
    assert np.array_equal(got_d, np.array([oracle.invsqrt2d(x) for x in xd]))
    # the published method's accuracy: ~4e-16 (double), ~1e-6 (float) relative
    assert np.abs(got_d * np.sqrt(xd) - 1.0).max() <= 5e-16
    assert np.abs(got_f.astype(np.float64) * np.sqrt(xf.astype(np.float64)) - 1.0).max() <= 5e-6
    t = [float(v) for v in lines[20000].split()[1:]]
    assert t_before - 1.0 <= t[0] <= t[1] <= time.time() + 1.0
    assert float(lines[20001].split()[1]) == pytest.approx(abs(-0.2475 + 0.25) / 0.25, rel=1e-15)


@pytest.mark.gpu
def test_debug_build_checks_every_evaluation():
    app = os.path.join(PKG, "bin", "nbody_app_debug")
    if not os.path.exists(app):
        pytest.fail(f"{app} is not built")
    r = subprocess.run([app, "--N", "64", "--Ncycle", "2", "--ic", "plummer", "--seed", "3", "--eps", "0.01",
                        "--dt", "0.001", "--energy", "physical"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    worst = [float(v) for l in r.stdout.splitlines() if l.startswith("max ") for v in l.split()[2:]]
    rows = [l for l in r.stdout.splitlines() if re.match(r"^[0-9.e+-]+\s+[0-9.e+-]+\s+[0-9.e+-]+$", l)]
    print(f"[gpu-debug] {len(worst) // 3} checks, worst per-component relative difference {max(worst):.2e}",
          flush=True)
    assert len(worst) == 3 * 2 * 3  # acc and adot, initial force + 2 steps
    assert len(rows) == 5 * 2 * 3   # the first PRT = 5 bodies printed each time
    assert max(worst) <= 1e-9
