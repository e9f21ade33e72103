"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Run in a checkout where the reference sources are present (oracle/Makefile
builds oracle/_ref/libref{64,32}.so from them):

    make -C oracle all && python tests/golden/make_golden.py

The reference hard-codes its softening (EPS2 = 1e-7, include/kernel.h:13),
so every vector here is at eps2 = 1e-7.  Initial conditions are the
reference's own NBody::initialize (src/nbody.cpp:69-114) or Plummer spheres
from the product's generator (stored in the fixture, so the tests do not
depend on the generator staying the same).  The reference runs on one
OpenMP thread, its only reproducible order.

Fixtures (numpy .npz, no pickles):
  ref_ic_n1024.npz   reference IC; FP64 and _F32 force_calculation on it;
                     compute_kinetic / compute_potential (E0 as printed)
  plummer_n1024.npz  Plummer seed 1: FP64 and _F32 force; 8 fixed-dt steps
                     (dt = 1/512) and 8 adaptive steps of NBody::timestepping
  tiny.npz           N = 2, 3, 8, 64: 4 fixed-dt steps, state and force
  plummer_n102400_sample.npz
                     Plummer seed 2, N = 102400: FP64 force of every 100th
                     body (full evaluation, subsampled) + IC checksum
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "n-body-code_amd"))

from oracle import Reference  # noqa: E402
import nbody_amd  # noqa: E402  (host-only IC generator; no device needed)


def state_sha(s):
    return hashlib.sha256(np.ascontiguousarray(s, dtype="<f8").tobytes()).hexdigest()


def main():
    r64 = Reference("fp64")
    r32 = Reference("f32")
    assert r64.eps2 == 1e-7

    ic = r64.initialize(1024)
    np.savez_compressed(os.path.join(HERE, "ref_ic_n1024.npz"),
                        state=ic, force64=r64.force(ic), force32=r32.force(ic),
                        kinetic=r64.kinetic(ic), potential=r64.potential(ic))

    p = nbody_amd.plummer(1024, 1)
    fixed = r64.run(p, 8, 1.0 / 512)
    adapt = r64.run(p, 8, 0.0)
    np.savez_compressed(os.path.join(HERE, "plummer_n1024.npz"),
                        state=p, force64=r64.force(p), force32=r32.force(p),
                        fixed_dt=1.0 / 512, fixed_state=fixed["state"], fixed_force=fixed["force"],
                        fixed_kin=fixed["kin"], fixed_pot=fixed["pot"],
                        adapt_state=adapt["state"], adapt_force=adapt["force"], adapt_dts=adapt["dts"])

    tiny = {}
    for n in (2, 3, 8, 64):
        s = nbody_amd.plummer(n, 100 + n)
        out = r64.run(s, 4, 1.0 / 256)
        tiny[f"state{n}"] = s
        tiny[f"final{n}"] = out["state"]
        tiny[f"force{n}"] = out["force"]
    np.savez_compressed(os.path.join(HERE, "tiny.npz"), dt=1.0 / 256, **tiny)

    t = time.time()
    big = nbody_amd.plummer(102400, 2)
    f = r64.force(big)
    idx = np.arange(0, 102400, 100)
    np.savez_compressed(os.path.join(HERE, "plummer_n102400_sample.npz"),
                        n=102400, seed=2, idx=idx, force64=f[:, idx], ic_sha256=state_sha(big))
    print(f"N=102400 reference force in {time.time() - t:.1f} s")


if __name__ == "__main__":
    main()
