"""Energy conservation at the benchmark's scale: the N = 1,048,576 Plummer
sphere (seed 3, eps = 1e-3, dt = 1/2048), four resident Hermite steps in
FP64 (the drop-in's directed pairing) and in mixed precision (symmetric
pairing, the benchmark's headline kernel), energy in FP64 on the device.

Bars: FP64 relative drift |E(T) - E(0)| / |E(0)| <= 1e-9 (it is dominated by
the integrator's truncation at this dt, ~3e-10); mixed drift within 10x of
FP64's (the target's criterion), and -- since both drifts are dominated by
the same truncation -- the mixed run's energy change within 1e-11 |E(0)| of
FP64's, which bounds what the mixed force error adds; final positions of
the two runs within 1e-9 of each other."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 1 << 20


def run(nb, precision, pairing, state):
    print(f"[gpu-drift] N={N} {precision} {pairing}: 4 steps", flush=True)
    with nb.Backend(precision, 1e-6, pairing=pairing) as b:
        b.upload(state)
        b.set_dt(1.0 / 2048)
        e0 = sum(b.energy())
        b.step(4, graph=True)
        e1 = sum(b.energy())
        s, _ = b.download()
    return e0, e1, s


def test_energy_drift_million_bodies(nb):
    state = nb.plummer(N, 3)
    e0, e1, s64 = run(nb, "fp64", "directed", state)
    m0, m1, smx = run(nb, "mixed", "symmetric", state)
    d64 = abs(e1 - e0) / abs(e0)
    dmx = abs(m1 - m0) / abs(m0)
    ddiff = abs((m1 - m0) - (e1 - e0)) / abs(e0)
    dpos = np.abs(smx[0:3] - s64[0:3]).max()
    print(f"[gpu-drift] E0 {e0:.12f}; drift fp64 {d64:.4e}, mixed {dmx:.4e}; "
          f"energy-change difference {ddiff:.2e}; max |x_mixed - x_fp64| {dpos:.2e}", flush=True)
    assert e0 == m0  # same state, same FP64 energy kernel
    assert d64 <= 1e-9
    assert dmx <= 10 * d64
    assert ddiff <= 1e-11
    assert dpos <= 1e-9
