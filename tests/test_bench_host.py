"""Host logic of bench.py, no GPU: the one-GPU refusal, the performed-flop
count of the symmetric legs against the block schedule the symmetric kernel
runs, kernel naming, and the traffic figures read from the committed PMC
summary."""
import importlib.util
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def schedule_pairs(n, b):
    """Unordered pairs between distinct blocks, summed over the symmetric
    kernel's schedule: full block P (of nr = n // b) takes blocks P+1 .. P+H
    (mod nr), H = (nr - 1) // 2, plus P + nr/2 when nr is even and P < nr/2,
    then the partial block of n - nr*b bodies, if any."""
    nr, part = n // b, n % b
    h = (nr - 1) // 2
    total = 0
    for p in range(nr):
        q = h + (1 if nr % 2 == 0 and p < nr // 2 else 0)
        total += q * b * b + b * part
    return total


@pytest.mark.parametrize("n,b", [(1 << 20, 768), (1 << 20, 512), (10 * 768, 768), (11 * 768 + 5, 768)])
def test_schedule_covers_every_cross_block_pair_once(n, b):
    # brute force over block labels: every unordered pair of bodies in
    # different blocks, counted once
    sizes = [b] * (n // b) + ([n % b] if n % b else [])
    brute = sum(sizes[i] * sizes[j] for i in range(len(sizes)) for j in range(i + 1, len(sizes)))
    assert schedule_pairs(n, b) == brute


@pytest.mark.parametrize("block,ipl", [(256, 3), (256, 2), (128, 3)])
def test_symmetric_performed_flop_matches_the_schedule(bench, block, ipl):
    n, b = bench.N, block * ipl
    shape = {"pairing": "symmetric", "block": block, "ipl": ipl}
    cross = schedule_pairs(n, b)
    # the block-diagonal pass: every ordered pair inside a block, self pairs excluded
    sizes = [b] * (n // b) + ([n % b] if n % b else [])
    inside = sum(s * (s - 1) for s in sizes)
    assert 2 * cross + inside + n == n * n
    assert bench.executed_flop(shape) == 73 * cross + 60 * inside
    assert bench.convention_flop(shape) == 60 * n * n


def test_directed_performed_flop(bench):
    assert bench.executed_flop({"pairing": "directed", "block": 256, "ipl": 1}) == 60 * bench.N ** 2


def test_kernel_names(bench):
    sym = {"shape": {"pairing": "symmetric", "block": 256, "ipl": 3}}
    assert bench.kname("mixed", sym) == "k_force_sym<SymMixed,3,4>"
    assert bench.kname("fp64", sym) == "k_force_sym<SymF64,3,4>"
    assert bench.kname("fp64", {"shape": {"pairing": "directed", "block": 128, "ipl": 2, "sources": "scalar"}}) \
        == "k_force_f64<128,2>"
    assert bench.kname("mixed", {"shape": {"pairing": "directed", "block": 64, "ipl": 1, "sources": "lds"}}) \
        == "k_force_mixed_lds<64,1>"


def test_traffic_read_from_the_committed_profile(bench):
    traffic, lease = bench.measured_traffic()
    assert lease and "commit" in lease
    for k in ("k_force_sym<SymMixed,3,4>", "k_force_sym<SymF64,3,4>"):
        # at least the compulsory 104 bytes per body
        assert traffic[k] >= 104 * bench.N
    # another block size of the same kernel and targets per lane stands in,
    # and says which instance was measured
    directed = [k for k in traffic if k.startswith("k_force_mixed<")]
    assert directed
    head, tail = directed[0].split("<", 1)
    ipl = tail.rstrip(">").split(",")[-1]
    other = f"{head}<{999},{ipl}>"
    val, inst = bench.traffic_of(traffic, other)
    assert inst in traffic and val == traffic[inst]
    assert bench.traffic_of(traffic, "k_force_nonexistent<64,1>") == (None, None)


def test_roofline_fraction_is_a_utilisation(bench):
    res = {"shape": {"pairing": "symmetric", "block": 256, "ipl": 3}, "executed_tflops": 90.0,
           "effective_tflops": 150.0, "kernel_ms": 440.0, "executed_flop": 4.0e13}
    r = bench.roofline("mixed", res, {}, None)
    assert r["peak"] == 157.3 and r["unit"] == "TFLOP/s"
    assert r["frac"] == pytest.approx(90.0 / 157.3)
    assert r["traffic"] is None and r["traffic_source"] is None
    r64 = bench.roofline("fp64", dict(res, executed_tflops=60.0), {}, None)
    assert r64["peak"] == 78.6 and r64["frac"] == pytest.approx(60.0 / 78.6)
