"""ctypes handles onto the oracle (TEST INFRASTRUCTURE ONLY).

`Oracle` wraps oracle/liboracle.so, the C restatement of the reference's
force/integrator path (oracle/oracle.c).  `Reference` wraps
oracle/_ref/libref{64,32}.so, the reference's own sources compiled by
oracle/Makefile, when they have been built.  Only tests/, the smoke() entry
point and bench.py's cpu_baseline leg import this module; the product never
does.

Arrays are numpy float64, contiguous.  A body state is a (7, n) array
(x, y, z, m, vx, vy, vz) -- the order of the reference's force_calculation
arguments, include/kernel.h:15-19 -- and forces are (6, n)
(ax, ay, az, jx, jy, jz).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def _p(a):
    return a.ctypes.data_as(_D)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """C restatement of the reference path."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        self.lib = L
        L.or_invsqrt2d.restype = C.c_double
        L.or_invsqrt2d.argtypes = [C.c_double]
        for name in ("or_force_f64", "or_force_f32"):
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = [C.c_int] + [_D] * 7 + [C.c_double] + [_D] * 6 + [C.c_int]
        L.or_force_sample_ld.restype = C.c_int
        L.or_force_sample_ld.argtypes = [C.c_int] + [_D] * 7 + [C.c_double, C.c_int, _I, _D, C.c_int]
        L.or_predict.restype = None
        L.or_predict.argtypes = [C.c_int, C.c_double] + [_D] * 12
        L.or_snap_correct.restype = None
        L.or_snap_correct.argtypes = [C.c_int, C.c_double] + [_D] * 12
        L.or_updatedt.restype = C.c_double
        L.or_updatedt.argtypes = [C.c_int, C.c_double] + [_D] * 4
        L.or_initialdt.restype = C.c_double
        L.or_initialdt.argtypes = [C.c_int, C.c_double, _D, _D]
        L.or_kinetic.restype = C.c_double
        L.or_kinetic.argtypes = [C.c_int] + [_D] * 4
        L.or_potential_compat.restype = C.c_double
        L.or_potential_compat.argtypes = [C.c_int] + [_D] * 4 + [C.c_double]
        L.or_potential.restype = C.c_double
        L.or_potential.argtypes = [C.c_int] + [_D] * 4 + [C.c_double, C.c_int]
        L.or_reference_ic.restype = C.c_int
        L.or_reference_ic.argtypes = [C.c_int] + [C.c_double] * 4 + [_D] * 7

    def invsqrt2d(self, x):
        return self.lib.or_invsqrt2d(float(x))

    def force(self, state, eps2, precision="fp64", nthreads=1):
        """Reference-order force (src/kernel.cpp:1164-1224 / :1101-1163)."""
        s = _f64(state)
        n = s.shape[1]
        out = np.zeros((6, n))
        f = self.lib.or_force_f64 if precision == "fp64" else self.lib.or_force_f32
        rc = f(n, *[_p(s[k]) for k in range(7)], float(eps2), *[_p(out[k]) for k in range(6)], int(nthreads))
        if rc != 0:
            raise MemoryError("oracle force: allocation failed")
        return out

    def force_sample(self, state, eps2, idx, nthreads=1):
        """Directed long-double acc/jerk of bodies idx from all others; (6, len(idx))."""
        s = _f64(state)
        n = s.shape[1]
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        out = np.zeros((len(idx), 6))
        self.lib.or_force_sample_ld(n, *[_p(s[k]) for k in range(7)], float(eps2), len(idx),
                                    idx.ctypes.data_as(_I), _p(out), int(nthreads))
        return out.T.copy()

    def predict(self, state, force, dt):
        """Predictor of src/nbody.cpp:152-163; returns the predicted state."""
        s = _f64(state).copy()
        f = _f64(force)
        self.lib.or_predict(s.shape[1], float(dt), _p(s[0]), _p(s[1]), _p(s[2]),
                            _p(s[4]), _p(s[5]), _p(s[6]),
                            *[_p(f[k]) for k in range(6)])
        return s

    def snap_correct(self, state_pred, f0, f1, dt):
        """Snap/crackle + corrector, src/nbody.cpp:165-196.

        Returns (corrected state, a2 (3,n), a3 (3,n))."""
        s = _f64(state_pred).copy()
        n = s.shape[1]
        a0 = _f64(f0[0:3]); j0 = _f64(f0[3:6])
        a1 = _f64(f1[0:3]); j1 = _f64(f1[3:6])
        a2 = np.zeros((3, n)); a3 = np.zeros((3, n))
        self.lib.or_snap_correct(n, float(dt), _p(s[0]), _p(s[1]), _p(s[2]), _p(s[4]), _p(s[5]), _p(s[6]),
                                 _p(a0), _p(j0), _p(a1), _p(j1), _p(a2), _p(a3))
        return s, a2, a3

    def updatedt(self, f0, a2, a3, eta=1e-2):
        n = f0.shape[1]
        a0 = _f64(f0[0:3]); j0 = _f64(f0[3:6])
        return self.lib.or_updatedt(n, float(eta), _p(a0), _p(j0), _p(_f64(a2)), _p(_f64(a3)))

    def initialdt(self, f0, eta0=1e-4):
        n = f0.shape[1]
        return self.lib.or_initialdt(n, float(eta0), _p(_f64(f0[0:3])), _p(_f64(f0[3:6])))

    def kinetic(self, state):
        s = _f64(state)
        return self.lib.or_kinetic(s.shape[1], _p(s[3]), _p(s[4]), _p(s[5]), _p(s[6]))

    def potential_compat(self, state, eps2=1e-7):
        s = _f64(state)
        return self.lib.or_potential_compat(s.shape[1], _p(s[0]), _p(s[1]), _p(s[2]), _p(s[3]), float(eps2))

    def potential(self, state, eps2, nthreads=1):
        s = _f64(state)
        return self.lib.or_potential(s.shape[1], _p(s[0]), _p(s[1]), _p(s[2]), _p(s[3]), float(eps2), int(nthreads))

    def reference_ic(self, n, vmax=1.0, rmax=1.0, mmin=0.01, mmax=1.0):
        s = np.zeros((7, n))
        rc = self.lib.or_reference_ic(n, vmax, rmax, mmin, mmax, *[_p(s[k]) for k in range(7)])
        if rc != 0:
            raise MemoryError("oracle reference_ic: allocation failed")
        return s

    def run(self, state, eps2, dt, nsteps, precision="fp64", nthreads=1, adaptive=False, dt0=None):
        """Reference integrator order (src/nbody.cpp:224-336): initial force,
        then per step predict -> force -> snap/correct -> (updatedt) -> swap.

        With adaptive=True the first step uses dt0 if given, else initialdt.
        Returns dict with the final state and force, the dts used and the
        next dt."""
        s = _f64(state).copy()
        f0 = self.force(s, eps2, precision, nthreads)
        cur = (dt0 if dt0 is not None else self.initialdt(f0)) if adaptive else dt
        dts = []
        for _ in range(nsteps):
            h = cur if adaptive else dt
            dts.append(h)
            p = self.predict(s, f0, h)
            f1 = self.force(p, eps2, precision, nthreads)
            s, a2, a3 = self.snap_correct(p, f0, f1, h)
            if adaptive:
                cur = self.updatedt(f0, a2, a3)
            f0 = f1
        return {"state": s, "force": f0, "dts": np.array(dts), "next_dt": cur}


class Reference:
    """The reference's own sources compiled into oracle/_ref (EPS2 = 1e-7)."""

    def __init__(self, precision="fp64"):
        name = "libref64.so" if precision == "fp64" else "libref32.so"
        path = os.path.join(HERE, "_ref", name)
        if not os.path.exists(path):
            raise RuntimeError(f"reference build missing: {path}")
        L = C.CDLL(path)
        self.lib = L
        L.ref_invsqrt2d.restype = C.c_double
        L.ref_invsqrt2d.argtypes = [C.c_double]
        L.ref_invsqrt2.restype = C.c_float
        L.ref_invsqrt2.argtypes = [C.c_float]
        L.ref_eps2.restype = C.c_double
        L.ref_force.restype = None
        L.ref_force.argtypes = [C.c_int] + [_D] * 13
        L.ref_force_threads.restype = None
        L.ref_force_threads.argtypes = [C.c_int, C.c_int] + [_D] * 13
        L.ref_kinetic.restype = C.c_double
        L.ref_kinetic.argtypes = [C.c_int] + [_D] * 4
        L.ref_potential.restype = C.c_double
        L.ref_potential.argtypes = [C.c_int] + [_D] * 4
        L.ref_initialize.restype = C.c_int
        L.ref_initialize.argtypes = [C.c_int] + [C.c_double] * 4 + [_D] * 7
        L.ref_run.restype = C.c_int
        L.ref_run.argtypes = [C.c_int, _D, C.c_int, C.c_double, _D, _D, _D, _D]

    @property
    def eps2(self):
        return self.lib.ref_eps2()

    def invsqrt2d(self, x):
        return self.lib.ref_invsqrt2d(float(x))

    def invsqrt2(self, x):
        return self.lib.ref_invsqrt2(float(x))

    def force(self, state, nthreads=1):
        """force_calculation of the compiled reference; nthreads > 1 runs its
        OpenMP loop threaded (results then depend on the merge order)."""
        s = _f64(state).copy()
        n = s.shape[1]
        out = np.zeros((6, n))
        if nthreads == 1:
            self.lib.ref_force(n, *[_p(s[k]) for k in range(7)], *[_p(out[k]) for k in range(6)])
        else:
            self.lib.ref_force_threads(n, int(nthreads), *[_p(s[k]) for k in range(7)],
                                       *[_p(out[k]) for k in range(6)])
        return out

    def kinetic(self, state):
        s = _f64(state).copy()
        return self.lib.ref_kinetic(s.shape[1], _p(s[3]), _p(s[4]), _p(s[5]), _p(s[6]))

    def potential(self, state):
        s = _f64(state).copy()
        return self.lib.ref_potential(s.shape[1], _p(s[0]), _p(s[1]), _p(s[2]), _p(s[3]))

    def initialize(self, n, vmax=1.0, rmax=1.0, mmin=0.01, mmax=1.0):
        s = np.zeros((7, n))
        self.lib.ref_initialize(n, vmax, rmax, mmin, mmax, *[_p(s[k]) for k in range(7)])
        return s

    def run(self, state, nsteps, dt_fixed=0.0):
        s = _f64(state).copy().reshape(-1)
        n = state.shape[1]
        force = np.zeros(6 * n)
        dts = np.zeros(nsteps + 1)
        kin = np.zeros(nsteps + 1)
        pot = np.zeros(nsteps + 1)
        self.lib.ref_run(n, _p(s), nsteps, float(dt_fixed), _p(force), _p(dts), _p(kin), _p(pot))
        return {"state": s.reshape(7, n), "force": force.reshape(6, n), "dts": dts, "kin": kin, "pot": pot}
