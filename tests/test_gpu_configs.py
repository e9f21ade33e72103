"""The benchmark configurations end to end: the mixed form's per-body error
against FP64 and its relative energy drift beside the FP64 drift.

  N=1024 Plummer seed 1, eps=1/64, dt=1/512, 256 steps
  N=102400 Plummer seed 2, eps=1/256, dt=1/1024, 3 steps
(N=1048576 is measured by bench.py; the 8192-cluster ensemble by
test_gpu_ensemble.py and bench.py.)"""
import numpy as np
import pytest

from conftest import body_relerr

pytestmark = pytest.mark.gpu


def _banner(msg):
    print(f"[gpu-configs] {msg}", flush=True)


def run(nb, state, prec, eps2, dt, nsteps):
    with nb.Backend(prec, eps2) as b:
        b.upload(state)
        f0 = b.download()[1]
        e0 = sum(b.energy())
        b.set_dt(dt)
        b.step(nsteps, graph=True)
        e1 = sum(b.energy())
        st, f1 = b.download()
    return f0, abs(e1 - e0) / abs(e0), st, f1


@pytest.mark.parametrize("cfg", [
    dict(name="N=1024", n=1024, seed=1, eps=1 / 64, dt=1 / 512, steps=256, sample=None),
    dict(name="N=102400", n=102400, seed=2, eps=1 / 256, dt=1 / 1024, steps=3, sample=384),
])
def test_mixed_vs_fp64(nb, oracle, cfg):
    s = nb.plummer(cfg["n"], cfg["seed"])
    eps2 = cfg["eps"] ** 2
    f64, d64, st64, _ = run(nb, s, "fp64", eps2, cfg["dt"], cfg["steps"])
    fmx, dmx, stmx, fmx1 = run(nb, s, "mixed", eps2, cfg["dt"], cfg["steps"])
    # per-body errors of both GPU paths against the CPU FP64 oracle on the
    # initial state (the whole system, or a random sample of bodies)
    if cfg["sample"] is None:
        ref = oracle.force(s, eps2)
        idx = slice(None)
    else:
        idx = np.sort(np.random.default_rng(1).choice(cfg["n"], cfg["sample"], replace=False))
        ref = oracle.force_sample(s, eps2, idx, nthreads=16)
    e64a = body_relerr(f64[:3, idx], ref[:3])
    e64j = body_relerr(f64[3:, idx], ref[3:])
    ema = body_relerr(fmx[:3, idx], ref[:3])
    emj = body_relerr(fmx[3:, idx], ref[3:])
    _banner(f"{cfg['name']}: fp64 acc max {e64a.max():.2e} jerk max {e64j.max():.2e} | "
            f"mixed acc median {np.median(ema):.2e} max {ema.max():.2e}, jerk median {np.median(emj):.2e} "
            f"max {emj.max():.2e} | energy drift over {cfg['steps']} steps fp64 {d64:.2e} mixed {dmx:.2e}")
    assert e64a.max() <= 1e-12 and e64j.max() <= 1e-12
    assert np.median(ema) <= 1e-6
    assert dmx <= max(10 * d64, 1e-10)
