#!/usr/bin/env python3
"""Benchmark: pair-interactions/s of the Hermite-4 step on one MI355X.

Workload: N = 1,048,576 equal-mass Plummer sphere (seed 3) in Henon units,
eps = 1e-3, dt = 1/2048, resident in HBM.  A step is one shared-timestep
Hermite step (predict, pair force + jerk over all N^2 directed pairs,
correct), run by the device-resident integrator.  Four legs run:
  fp64           -- the drop-in (plain FP64 pair arithmetic, directed pairs,
                    bit-reproducible);
  fp64_symmetric -- FP64, each unordered pair evaluated once and applied to
                    both bodies (symmetric pairing, the reference's j < i
                    scheme), totals combined with FP64 atomics;
  mixed          -- FP64 position differences rounded to FP32, FP32 r^2 /
                    rsqrt / jerk, two-sum compensated FP32 sums, symmetric
                    pairing: the headline `value`;
  mixed_directed -- the mixed arithmetic on directed pairs.
`value` is the mixed leg's N^2 * steps / time.  It is valid only where its
median per-body relative acceleration error against FP64 is <= 1e-6; that
error is measured in the same run on the same state.  mixed_over_fp64 gives
the like-for-like ratios (symmetric/symmetric, directed/directed) and the
symmetric mixed leg over the directed FP64 drop-in.

Roofline: the pair kernels are VALU-bound (no MFMA; ~1e4-1e5 flop per HBM
byte).  Algorithmic work, for either pairing: 60 flop per directed pair (acc +
jerk convention) over the N^2 directed pairs of an evaluation, the count used
for the reference's own triangle loop as well.  achieved = that work / the
kernels' mean launch time from HIP events recorded on their own stream in the
timed region; peak = MI355X vector peak (FP32 157.3 TFLOP/s for mixed, FP64
78.6 TFLOP/s for FP64).  For the symmetric kernels, which evaluate each
unordered pair once, `executed` restates the rate with the work they do: 73
flop per unordered pair (the 60 plus m_i r^3 and six multiply-adds for the
reaction) and 60 per directed pair of the block-diagonal pass.

cpu_baseline: the oracle's restatement of the reference FP64 force loop
(src/kernel.cpp:1164-1224: Newton-3rd-law triangle, cyclic rows per thread,
per-thread sums merged at the end) on the first M bodies of the same
sphere, OpenMP threads on this host's cores, M sized for ~10-20 s.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "n-body-code_amd"))

N = 1 << 20
SEED = 3
EPS = 1e-3
DT = 1.0 / 2048
FLOP_PER_PAIR = 60
FLOP_PER_SYM_PAIR = 73
PEAK = {"fp64": 78.6, "mixed": 157.3}  # TFLOP/s, MI355X vector
SYM_FLUSH_TILES = {"fp64": 8, "mixed": 2}  # tiles per flush of the symmetric kernels' target sums


def body_relerr(a, b):
    return np.sqrt(((a - b) ** 2).sum(0)) / np.sqrt((b ** 2).sum(0))


def calibration_copy(nbytes=256 << 20):
    """One device-to-device copy of a known size (a runtime blit kernel), so
    that rocprofv3's FETCH_SIZE / WRITE_SIZE per dispatch can be scaled to
    bytes: run under the profiler with --pmc-calibrate."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    try:
        assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
        assert hip.hipDeviceSynchronize() == 0
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def measured_traffic():
    """HBM bytes per launch of the pair kernels, as measured by rocprofv3
    --pmc FETCH_SIZE / WRITE_SIZE passes (scaled by the calibration copy) and
    recorded in profiles/pmc_force_n1048576.md; None when not recorded."""
    path = os.path.join(ROOT, "profiles", "pmc_force_n1048576.md")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch "):
                _, kern, val = line.split()
                out[kern] = float(val)
    return out


def kname(prec, res):
    """Pair-kernel instance of a run: k_force_{f64,mixed}[_lds]<block,targets per
    lane> (scalar-load or LDS-tile sources) or the symmetric kernel
    k_force_sym<SymF64|SymMixed,targets per lane,waves>."""
    sh = res["shape"]
    if sh.get("pairing") == "symmetric":
        return f"k_force_sym<{'SymF64' if prec == 'fp64' else 'SymMixed'},{sh['ipl']},{sh['block'] // 64}>"
    lds = "_lds" if sh.get("sources") == "lds" else ""
    return f"k_force_{'f64' if prec == 'fp64' else 'mixed'}{lds}<{sh['block']},{sh['ipl']}>"


def kernel_flop(shape):
    """Algorithmic flop of one force evaluation: 60 per directed pair, N^2 pairs."""
    return FLOP_PER_PAIR * N * N


def sym_block(shape):
    """Bodies per block of the symmetric kernel: threads x targets per lane."""
    return shape["block"] * shape["ipl"]


def executed_flop(shape):
    """Flop the launch shape's kernels perform for one evaluation."""
    if shape.get("pairing") == "symmetric":
        b = sym_block(shape)
        inside = (N // b) * b * b + (N % b) ** 2      # ordered pairs inside blocks, self pairs included
        return FLOP_PER_SYM_PAIR * (N * N - inside) // 2 + FLOP_PER_PAIR * (inside - N)
    return FLOP_PER_PAIR * N * N


def run_precision(nb, prec, state, steps, warmup, log, pairing="directed"):
    tag = prec if pairing == "directed" or prec == "fp64" else f"{prec}/{pairing}"
    with nb.Backend(prec, EPS * EPS, pairing=pairing) as b:
        tune = None
        if b.launch_shape(N)["pairing"] == "directed":
            tune = b.autotune(N)        # launch shape for this size bucket, timed on the device
        log(f"{tag}: shape {b.launch_shape(N)}, autotune {tune}")
        b.upload(state)                 # initial force, NBody::cycle
        f0 = b.download()[1]
        b.set_dt(DT)
        if warmup:
            b.step(warmup)
        e0 = sum(b.energy())
        b.reset_timing()
        b.set_profiling(True)
        t0 = time.perf_counter()
        b.step(steps)                   # synchronous: stream drained on return
        t1 = time.perf_counter()
        b.set_profiling(False)
        tm = b.timing()
        e1 = sum(b.energy())
        shape = b.launch_shape(N)
        # drop-in boundary: host arrays in and out (PCIe included), as
        # force_calculation is called
        t2 = time.perf_counter()
        _, dropin_kernel_ms = b.force(state, return_ms=True)
        dropin_wall = time.perf_counter() - t2
    sec = t1 - t0
    kms = tm["force_ms"] / max(tm["force_launches"], 1)
    flop = kernel_flop(shape)
    xflop = executed_flop(shape)
    res = {
        "pairs_per_s": N * N * steps / sec,
        "ms_per_step": 1e3 * sec / steps,
        "steps_per_s": steps / sec,
        "kernel_ms": kms,
        "kernel_launches": tm["force_launches"],
        "kernel_flop": flop,
        "kernel_tflops": flop / (kms * 1e-3) / 1e12,
        "executed_flop": xflop,
        "executed_tflops": xflop / (kms * 1e-3) / 1e12,
        "energy_drift": abs(e1 - e0) / abs(e0),
        "shape": shape,
        "autotune": tune,
        "dropin": {"wall_ms": 1e3 * dropin_wall, "kernel_ms": dropin_kernel_ms,
                   "pairs_per_s_pcie_inclusive": N * N / dropin_wall},
    }
    log(f"{tag}: {res['ms_per_step']:.1f} ms/step, kernel {kms:.1f} ms, "
        f"{res['pairs_per_s']:.3e} pairs/s, drift {res['energy_drift']:.2e}")
    return res, f0


def traffic_of(traffic, kern):
    """Measured HBM bytes per launch of `kern`, or of the same kernel with
    another block size (the autotuner may pick either from run to run);
    returns (bytes, instance measured) or (None, None)."""
    if kern in traffic:
        return traffic[kern], kern
    head, tail = kern.split("<", 1)
    ipl = tail.rstrip(">").split(",")[-1]
    for k, v in sorted(traffic.items()):
        kh, kt = k.split("<", 1)
        if kh == head and kt.rstrip(">").split(",")[-1] == ipl:
            return v, k
    return None, None


def roofline(prec, res, traffic):
    tb, tk = traffic_of(traffic, kname(prec, res))
    out = {"bound": "valu", "achieved": res["kernel_tflops"], "peak": PEAK[prec],
           "unit": "TFLOP/s", "frac": res["kernel_tflops"] / PEAK[prec],
           "bound_note": "vector-ALU FLOP roofline (MI355X vector peak: FP64 78.6, FP32 157.3 TFLOP/s): "
                         "the softened r^-3 pair sum is VALU work, HBM-bound by no measure and no MFMA contraction",
           "traffic": tb, "traffic_kernel": tk, "algorithmic_bytes": 104 * N,
           "kernel": kname(prec, res), "kernel_ms": res["kernel_ms"],
           "flop_per_launch": res["kernel_flop"],
           "flop_convention": "60 flop per directed pair, N^2 directed pairs per evaluation"}
    if res["shape"].get("pairing") == "symmetric":
        b = sym_block(res["shape"])
        fl = SYM_FLUSH_TILES[prec]
        out["traffic_note"] = (f"mostly FP64 atomic adds: the per-tile source totals, about 3 N^2 / {b} per launch "
                               f"(blocks of {b}), and the target sums, flushed every {fl} tiles of 64 sources "
                               f"(3 N^2 / {64 * fl}), performed at the memory side (the 8 XCDs' L2s are not "
                               "coherent); besides, each workgroup reads its target block and source blocks "
                               "through L2. algorithmic_bytes = 7 doubles read + 6 written per body")
        out["executed"] = {"flop_per_launch": res["executed_flop"], "achieved": res["executed_tflops"],
                           "frac": res["executed_tflops"] / PEAK[prec],
                           "convention": "73 flop per unordered pair + 60 per directed pair of the "
                                         "block-diagonal pass"}
    else:
        out["traffic_note"] = ("the source records re-read through L2 by every workgroup, and the j-split "
                               "partial planes written and reduced by a second kernel. "
                               "algorithmic_bytes = 7 doubles read + 6 written per body")
    return out


def run_ensemble(nb, log, nsteps=64):
    """8192 independent Plummer clusters, n_k log-uniform in [8, 4096]
    (seed 4), eps_k = 1/(4 n_k^(1/3)), dt = 1/256: autotuned persistent
    force launch, steps replayed from a hipGraph.  The *_symmetric legs
    evaluate systems of >= 1024 bodies with symmetric pairing (a second
    launch beside the persistent one)."""
    counts = nb.ensemble_sizes(8192, 8, 4096, 4)
    eps2 = (1.0 / (4.0 * counts.astype(float) ** (1.0 / 3.0))) ** 2
    state = nb.ensemble_state(counts, 4)
    pairs = float((counts.astype(np.float64) ** 2).sum())
    out = {"workload": "8192 Plummer clusters, n_k log-uniform in [8,4096] (seed 4), "
                       "eps_k=1/(4 n_k^(1/3)), dt=1/256, 64 steps, hipGraph replay",
           "n_systems": int(len(counts)), "n_bodies": int(counts.sum()), "pairs_per_step": pairs}
    for leg in ("mixed", "fp64", "mixed_symmetric", "fp64_symmetric"):
        prec = leg.split("_")[0]
        pairing = "symmetric" if leg.endswith("symmetric") else "directed"
        with nb.Backend(prec, pairing=pairing) as b, nb.Ensemble(b, counts, eps2) as e:
            e.upload(state)
            tune = e.autotune()
            e.set_dt(1.0 / 256)
            e.step(16, graph=True)      # warm-up: builds and replays the graph once
            k0, u0 = e.energy()
            t0 = time.perf_counter()
            e.step(nsteps, graph=True)
            sec = time.perf_counter() - t0
            k1, u1 = e.energy()
            info = e.info()
        drift = np.abs((k1 + u1) - (k0 + u0)) / np.abs(k0 + u0)
        out[leg] = {"systems_per_s": len(counts) * nsteps / sec, "pairs_per_s": pairs * nsteps / sec,
                     "steps_per_s": nsteps / sec, "ms_per_step": 1e3 * sec / nsteps,
                     "energy_drift_median": float(np.median(drift)), "energy_drift_max": float(drift.max()),
                     "tuning": tune, "work_items": info["items"], "workgroups": info["workgroups"]}
        log(f"ensemble {leg}: {1e3 * sec / nsteps:.2f} ms/step, {pairs * nsteps / sec:.3e} pairs/s, "
            f"{len(counts) * nsteps / sec:.3e} systems/s, drift median {np.median(drift):.1e}, tuning {tune}")
    return out


def run_adjoint(nb, log):
    """Differentiation workload: N=4096 Plummer (seed 5), masses x
    lognormal(0.1) renormalised to 1, eps=1/64, dt=1/512, 128 FP64 steps,
    checkpoints every 16; loss = sum |x(T) - x_target|^2 with targets from a
    run with re-drawn masses (seed 6).  Times the forward run and the
    checkpointed backward sweep (which recomputes each segment)."""
    n, dt, nsteps, every = 4096, 1.0 / 512, 128, 16
    s = nb.plummer(n, 5)
    m = s[3] * np.random.Generator(np.random.PCG64(5)).lognormal(0.0, 0.1, n)
    s[3] = m / m.sum()
    st = s.copy()
    mt = np.random.Generator(np.random.PCG64(6)).lognormal(0.0, 0.1, n)
    st[3] = mt / mt.sum()
    with nb.Backend("fp64", (1.0 / 64) ** 2) as b, nb.Adjoint(b, n, every) as adj:
        target = adj.forward(st, dt, nsteps)[0:3]
        t0 = time.perf_counter()
        fin = adj.forward(s, dt, nsteps)
        t1 = time.perf_counter()
        cot = np.zeros((7, n))
        cot[0:3] = 2.0 * (fin[0:3] - target)
        g = adj.backward(cot)
        t2 = time.perf_counter()
    out = {"workload": "N=4096 Plummer seed 5, lognormal(0.1) masses, eps=1/64, dt=1/512, 128 FP64 steps, "
                       "checkpoints every 16",
           "forward_s": t1 - t0, "backward_s": t2 - t1,
           "loss": float(((fin[0:3] - target) ** 2).sum()),
           "grad_norm": {"x": float(np.linalg.norm(g[0:3])), "v": float(np.linalg.norm(g[4:7])),
                         "m": float(np.linalg.norm(g[3]))}}
    log(f"adjoint: forward {out['forward_s'] * 1e3:.1f} ms, backward {out['backward_s'] * 1e3:.1f} ms")
    return out


def _time_force(run, state, budget_s):
    """One evaluation of run() on the first M bodies, M sized for ~budget_s."""
    m0 = 16384
    t = time.perf_counter()
    run(state[:, :m0])
    rate = m0 * (m0 - 1) / (time.perf_counter() - t)
    m = int(np.sqrt(rate * budget_s))
    m = int(min(max(m, m0), 262144)) // 1024 * 1024
    t = time.perf_counter()
    run(state[:, :m])
    sec = time.perf_counter() - t
    return m, sec, m * (m - 1) / sec


def cpu_baseline(state, budget_s=12.0):
    """The reference's CPU force_calculation, FP64 (src/kernel.cpp:1164-1224)
    and _F32 (src/kernel.cpp:1101-1163) builds, compiled from its own sources
    into oracle/_ref (EPS2 = 1e-7: same work per pair), on the first M bodies
    of the same sphere with OpenMP threads on this host's cores; when that
    build is absent, the oracle's restatement of the same loops.  M is sized
    for ~budget_s of CPU time each."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, Reference
    cores = max(1, min(16, os.cpu_count() or 1))
    runs = {}
    try:
        for prec in ("fp64", "f32"):
            R = Reference(prec)
            runs[prec] = (lambda s, R=R: R.force(s, nthreads=cores))
        kind = "reference"
        what = "the reference's own {} force_calculation (compiled from its sources, EPS2=1e-7)"
    except (RuntimeError, OSError):
        O = Oracle()
        for prec in ("fp64", "f32"):
            runs[prec] = (lambda s, p=prec: O.force(s, EPS * EPS, p, nthreads=cores))
        kind = "port"
        what = "the oracle's restatement of the reference {} loop (eps=1e-3)"
    out = {}
    for prec in ("fp64", "f32"):
        m, sec, rate = _time_force(runs[prec], state, budget_s)
        out[prec] = (m, sec, rate)
    m, sec, rate = out["fp64"]
    m32, sec32, rate32 = out["f32"]
    return {"value": rate, "unit": "pair-interactions/s", "cores": cores, "kind": kind,
            "sample": f"one evaluation of {what.format('FP64')}, triangle order, over the first {m} bodies of "
                      f"the N=1048576 seed-3 Plummer sphere, {cores} OpenMP threads (one process), {sec:.1f} s; "
                      f"rate = M(M-1)/t directed pairs",
            "f32": {"value": rate32, "unit": "pair-interactions/s", "cores": cores, "kind": kind,
                    "sample": f"the same with {what.format('_F32')}, first {m32} bodies, {sec32:.1f} s"}}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--pmc-calibrate", action="store_true",
                    help="start with a 256 MiB device-to-device copy (profiler calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU split is design-only).", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    log = (lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)) if rank == 0 else (lambda m: None)

    import nbody_amd as nb
    log(f"device {nb.device_info()}")
    if args.pmc_calibrate:
        calibration_copy()
    t = time.perf_counter()
    state = nb.plummer(N, SEED)
    log(f"Plummer N={N} in {time.perf_counter() - t:.1f} s")

    r64, f64 = run_precision(nb, "fp64", state, args.steps, args.warmup, log)
    r64s, f64s = run_precision(nb, "fp64", state, args.steps, args.warmup, log, pairing="symmetric")
    rmx, fmx = run_precision(nb, "mixed", state, args.steps, args.warmup, log, pairing="symmetric")
    rmd, fmd = run_precision(nb, "mixed", state, args.steps, args.warmup, log)
    err = body_relerr(fmx[:3], f64[:3])
    med, mx = float(np.median(err)), float(err.max())
    errd = body_relerr(fmd[:3], f64[:3])
    # symmetric vs directed FP64, per body, acceleration and jerk
    err64s = np.maximum(body_relerr(f64s[:3], f64[:3]), body_relerr(f64s[3:], f64[3:]))
    log(f"mixed vs fp64 acc rel err: median {med:.2e}, max {mx:.2e} "
        f"(directed: median {np.median(errd):.2e}, max {errd.max():.2e}); "
        f"fp64 symmetric vs directed max {err64s.max():.2e}")

    ens = None if args.no_ensemble else run_ensemble(nb, log)
    adj = None if args.no_ensemble else run_adjoint(nb, log)

    cpu = None
    if not args.no_cpu_baseline and rank == 0:
        cpu = cpu_baseline(state)
        log(f"cpu baseline {cpu['value']:.3e} pairs/s (FP64), {cpu['f32']['value']:.3e} (_F32) "
            f"on {cpu['cores']} threads, {cpu['kind']}")

    traffic = measured_traffic()
    out = {
        "metric": "pair-interactions/s at N=1M: mixed vs FP64 GPU @ median rel acc err <=1e-6",
        "value": rmx["pairs_per_s"],
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rmx["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64/f32 mixed",
        "data": "synthetic (Plummer sphere, generated in-process)",
        "config": {"workload": "Hermite-4 shared step, N=1048576 Plummer seed 3, eps=1e-3, dt=1/2048",
                   "n_bodies": N, "global_batch": N, "seq_len": None, "parallelism": "none (1 GPU)"},
        "valid": med <= 1e-6,
        "mixed_median_rel_acc_err": med,
        "mixed_max_rel_acc_err": mx,
        "fp64_pairs_per_s": r64["pairs_per_s"],
        "mixed_over_fp64": {"symmetric_over_symmetric": rmx["pairs_per_s"] / r64s["pairs_per_s"],
                            "directed_over_directed": rmd["pairs_per_s"] / r64["pairs_per_s"],
                            "symmetric_over_directed_dropin": rmx["pairs_per_s"] / r64["pairs_per_s"]},
        "mixed_pairing": rmx["shape"]["pairing"],
        "fp64_symmetric": {"pairs_per_s": r64s["pairs_per_s"], "over_fp64_directed": r64s["pairs_per_s"] / r64["pairs_per_s"],
                           "max_rel_err_vs_directed": float(err64s.max()), "energy_drift": r64s["energy_drift"],
                           "launch_shape": r64s["shape"], "roofline": roofline("fp64", r64s, traffic)},
        "mixed_directed": {"pairs_per_s": rmd["pairs_per_s"], "over_fp64": rmd["pairs_per_s"] / r64["pairs_per_s"],
                           "median_rel_acc_err": float(np.median(errd)), "max_rel_acc_err": float(errd.max()),
                           "energy_drift": rmd["energy_drift"], "launch_shape": rmd["shape"],
                           "roofline": roofline("mixed", rmd, traffic)},
        "steps_per_s": {"fp64": r64["steps_per_s"], "fp64_symmetric": r64s["steps_per_s"],
                        "mixed": rmx["steps_per_s"], "mixed_directed": rmd["steps_per_s"]},
        "energy_drift": {"fp64": r64["energy_drift"], "fp64_symmetric": r64s["energy_drift"],
                         "mixed": rmx["energy_drift"], "mixed_directed": rmd["energy_drift"]},
        "launch_shape": rmx["shape"],
        "dropin_host_arrays": {"fp64": r64["dropin"], "mixed": rmx["dropin"]},
        "roofline": roofline("mixed", rmx, traffic),
        "roofline_fp64": roofline("fp64", r64, traffic),
        "cpu_baseline": cpu,
        "ensemble": ens,
        "adjoint": adj,
    }
    if rank == 0:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
