"""The drop-in proven with the reference's own integrator: its
src/nbody.cpp and src/utils.cpp, compiled where they lie with the `_HIP`
branch of force_calculation (oracle/ref_hip_force.cpp, the code
INTEGRATION.md shows) in place of its src/kernel.cpp, into
oracle/_ref/libref_hip.so.  NBody::timestepping (src/nbody.cpp:224-289)
itself then calls the HIP library for every force.

Eight fixed-dt steps (dt = 1/512) of the 1024-body Plummer sphere (seed 1)
at the reference's EPS2 = 1e-7 must reproduce the reference's own CPU run
of the same steps (plummer_n1024.npz fixed_state / fixed_force, written by
the compiled reference): FP64 to 1e-13 relative in state and 1e-12 per body
in force (the sums differ from the reference's triangle only in order);
mixed precision to the mixed kernels' error carried through eight steps.

The reference's NBody needs MPI (src/nbody.cpp:22-32), so the run happens
in a child process: an MPI start-up failure there cannot take the test
session down with it.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, body_relerr, golden

pytestmark = pytest.mark.gpu

LIB = os.path.join(ROOT, "oracle", "_ref", "libref_hip.so")

CHILD = r"""
import ctypes as C, sys
import numpy as np
lib, prec, pair, src, dst = sys.argv[1], int(sys.argv[2]), int(sys.argv[3]), sys.argv[4], sys.argv[5]
g = np.load(src, allow_pickle=False)
L = C.CDLL(lib)
D = C.POINTER(C.c_double)
L.refhip_run.restype = C.c_int
L.refhip_run.argtypes = [C.c_int, C.c_int, C.c_int, D, C.c_int, C.c_double, D]
state = np.ascontiguousarray(g["state"], dtype=np.float64).copy()
n = state.shape[1]
force = np.empty((6, n))
cores = L.refhip_run(n, prec, pair, state.ctypes.data_as(D), 8, float(g["fixed_dt"]), force.ctypes.data_as(D))
np.savez(dst, state=state, force=force, cores=cores)
print("refhip child done", flush=True)
"""


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_reference_timestepping_on_hip_forces(nb, tmp_path, precision):
    if not os.path.exists(LIB):
        pytest.skip(f"{LIB} is not built (it needs the reference sources at build time)")
    src = os.path.join(ROOT, "tests", "golden", "plummer_n1024.npz")
    dst = str(tmp_path / "out.npz")
    print(f"[gpu-refhip] reference NBody::timestepping x8 on HIP forces, {precision}", flush=True)
    env = dict(os.environ, MPIR_CVAR_CH3_INTERFACE_HOSTNAME="127.0.0.1")
    r = subprocess.run([sys.executable, "-c", CHILD, LIB, str(nb.PRECISIONS[precision]), "0", src, dst],
                       capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.load(dst, allow_pickle=False)
    state, force = out["state"], out["force"]
    assert int(out["cores"]) == nb.device_info()["compute_units"]  # Sys::num_cores filled by the branch
    g = golden("plummer_n1024.npz")
    ref, fref = g["fixed_state"], g["fixed_force"]
    assert np.array_equal(state[3], ref[3])
    dx = np.abs(state[0:3] - ref[0:3]).max() / np.abs(ref[0:3]).max()
    dv = np.abs(state[4:7] - ref[4:7]).max() / np.abs(ref[4:7]).max()
    ea = body_relerr(force[:3], fref[:3])
    ej = body_relerr(force[3:], fref[3:])
    print(f"[gpu-refhip]   state vs reference: x {dx:.2e}, v {dv:.2e}; force acc max {ea.max():.2e}, "
          f"jerk max {ej.max():.2e}", flush=True)
    if precision == "fp64":
        assert dx <= 1e-13 and dv <= 1e-13
        assert ea.max() <= 1e-12 and ej.max() <= 1e-12
    else:
        assert dx <= 1e-6 and dv <= 1e-6
        assert np.median(ea) <= 1e-6 and np.median(ej) <= 1e-5
