"""Python handle onto libnbody_hip.so (the C ABI of include/nbody_hip.h).

Thin ctypes layer for tests, the benchmark and Python callers.  Every call
goes to the HIP library; there is no CPU fallback: if the library is not
built, importing succeeds but any use raises `BackendMissing`.

Arrays are numpy float64.  A body state is a (7, n) array in the order of
the reference's force_calculation arguments (include/kernel.h:15-19):
x, y, z, m, vx, vy, vz.  Forces are (6, n): ax, ay, az, jx, jy, jz.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(ROOT, "lib", "libnbody_hip.so")
HEADER = os.path.join(os.path.dirname(ROOT), "include", "nbody_hip.h")

FP64 = 0
MIXED = 1
PRECISIONS = {"fp64": FP64, "mixed": MIXED}

OK, ERR_ARG, ERR_HIP, ERR_ALLOC, ERR_STATE, ERR_NONFINITE = range(6)

_D = C.POINTER(C.c_double)


class BackendMissing(RuntimeError):
    pass


class NBodyError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [("device", C.c_int), ("precision", C.c_int), ("eps2", C.c_double),
                ("block", C.c_int), ("jsplit", C.c_int), ("ipl", C.c_int), ("pairing", C.c_int),
                ("sources", C.c_int)]


PAIRINGS = {"directed": 0, "symmetric": 1}
SOURCES = {"auto": 0, "lds": 1, "scalar": 2}  # enum nbh_sources


_lib = None


def lib():
    """Loads the HIP library once; raises BackendMissing if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BackendMissing(f"{LIB_PATH} is not built (run __graft_entry__.build() or make -C n-body-code_amd)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.nbh_last_error.restype = C.c_char_p
    L.nbh_version.restype = C.c_int
    L.nbh_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
    L.nbh_destroy.argtypes = [vp]
    L.nbh_configure.argtypes = [vp, C.POINTER(Config)]
    L.nbh_force.argtypes = [vp, C.c_int] + [_D] * 13 + [_D]
    L.nbh_force_device.argtypes = [vp, C.c_int] + [vp] * 13 + [vp]
    L.nbh_force_device_noalloc.argtypes = [vp, C.c_int] + [vp] * 13 + [vp]
    L.nbh_upload.argtypes = [vp, C.c_int] + [_D] * 7
    L.nbh_initial_dt.argtypes = [vp, C.c_double, _D]
    L.nbh_set_dt.argtypes = [vp, C.c_double, C.c_int, C.c_double]
    L.nbh_get_dt.argtypes = [vp, _D]
    L.nbh_step.argtypes = [vp, C.c_int, C.c_int]
    L.nbh_download.argtypes = [vp] + [_D] * 13
    L.nbh_energy.argtypes = [vp, C.c_int, _D, _D]
    L.nbh_energy_host.argtypes = [vp, C.c_int] + [_D] * 7 + [C.c_int, _D, _D]
    L.nbh_set_profiling.argtypes = [vp, C.c_int]
    L.nbh_timing.argtypes = [vp, _D, C.POINTER(C.c_int64), _D, C.POINTER(C.c_int64)]
    L.nbh_reset_timing.argtypes = [vp]
    L.nbh_launch_shape.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.nbh_launch_ipl.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.nbh_launch_pairing.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.nbh_launch_sources.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.nbh_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                               C.POINTER(C.c_double)]
    L.nbh_device_info.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_char_p, C.c_int]
    L.nbh_plummer.argtypes = [C.c_int, C.c_uint64] + [_D] * 7
    _I = C.POINTER(C.c_int)
    L.nbh_ens_create.argtypes = [vp, C.c_int, _I, _D, C.POINTER(vp)]
    L.nbh_ens_destroy.argtypes = [vp]
    L.nbh_ens_upload.argtypes = [vp] + [_D] * 7
    L.nbh_ens_set_dt.argtypes = [vp, C.c_double]
    L.nbh_ens_step.argtypes = [vp, C.c_int, C.c_int]
    L.nbh_ens_download.argtypes = [vp] + [_D] * 13
    L.nbh_ens_energy.argtypes = [vp, _D, _D]
    L.nbh_ens_autotune.argtypes = [vp, _I, _I, _D]
    L.nbh_ens_set_tuning.argtypes = [vp, _I, _I]
    L.nbh_ens_set_targets_per_lane.argtypes = [vp, C.c_int]
    L.nbh_ens_tuning.argtypes = [vp, _I, _I, C.POINTER(C.c_int)]
    L.nbh_ens_info.argtypes = [vp, _I, _I, _I]
    L.nbh_reference_ic.argtypes = [C.c_int] + [C.c_double] * 4 + [_D] * 7
    L.nbh_force_vjp.argtypes = [vp, C.c_int, _D, _D, _D]
    L.nbh_state_save.argtypes = [C.c_char_p, C.c_int] + [_D] * 7 + [C.c_double, C.c_int64]
    L.nbh_state_info.argtypes = [C.c_char_p, C.POINTER(C.c_int), _D, C.POINTER(C.c_int64)]
    L.nbh_state_load.argtypes = [C.c_char_p, C.c_int] + [_D] * 7
    L.nbh_adj_create.argtypes = [vp, C.c_int, C.c_int, C.POINTER(vp)]
    L.nbh_adj_destroy.argtypes = [vp]
    L.nbh_adj_forward.argtypes = [vp, _D, C.c_double, C.c_int, _D]
    L.nbh_adj_backward.argtypes = [vp, _D, _D]
    _lib = L
    return L


def _check(rc):
    if rc != OK:
        raise NBodyError(rc, lib().nbh_last_error().decode())


def _p(a):
    return a.ctypes.data_as(_D)


def _state(state):
    s = np.ascontiguousarray(state, dtype=np.float64)
    if s.ndim != 2 or s.shape[0] != 7:
        raise ValueError(f"state must be (7, n), got {s.shape}")
    return s


def plummer(n, seed):
    """Equal-mass Plummer sphere in Henon units (host code); (7, n)."""
    s = np.empty((7, n))
    rc = lib().nbh_plummer(int(n), int(seed), *[_p(s[k]) for k in range(7)])
    if rc != OK:
        raise NBodyError(rc, f"nbh_plummer(n={n}) rejected its arguments")
    return s


def reference_ic(n, vmax=1.0, rmax=1.0, mmin=0.01, mmax=1.0):
    """NBody::initialize of the reference CLI (src/nbody.cpp:69-114); (7, n)."""
    s = np.empty((7, n))
    rc = lib().nbh_reference_ic(int(n), vmax, rmax, mmin, mmax, *[_p(s[k]) for k in range(7)])
    if rc != OK:
        raise NBodyError(rc, f"nbh_reference_ic(n={n}) rejected its arguments")
    return s


def state_save(path, state, time=0.0, step=0):
    """Writes a (7, n) state to a body-state file (nbh_state_save)."""
    s = _state(state)
    _check(lib().nbh_state_save(os.fsencode(path), s.shape[1], *[_p(s[k]) for k in range(7)],
                                float(time), int(step)))


def state_info(path):
    n, t, st = C.c_int(), C.c_double(), C.c_int64()
    _check(lib().nbh_state_info(os.fsencode(path), C.byref(n), C.byref(t), C.byref(st)))
    return {"n": n.value, "time": t.value, "step": st.value}


def state_load(path):
    """Reads a body-state file into a (7, n) array."""
    n = state_info(path)["n"]
    s = np.empty((7, n))
    _check(lib().nbh_state_load(os.fsencode(path), n, *[_p(s[k]) for k in range(7)]))
    return s


def device_info(device=0):
    cu, ws = C.c_int(), C.c_int()
    name = C.create_string_buffer(64)
    _check(lib().nbh_device_info(device, C.byref(cu), C.byref(ws), name, 64))
    return {"compute_units": cu.value, "wave_size": ws.value, "arch": name.value.decode()}


class Backend:
    """One nbh_ctx: a device, a stream, buffers, a resident state."""

    def __init__(self, precision="fp64", eps2=1e-7, device=0, block=0, jsplit=0, ipl=0, pairing="directed",
                 sources="auto"):
        L = lib()
        self._cfg = Config(device, PRECISIONS[precision] if isinstance(precision, str) else int(precision),
                           float(eps2), int(block), int(jsplit), int(ipl),
                           PAIRINGS[pairing] if isinstance(pairing, str) else int(pairing),
                           SOURCES[sources] if isinstance(sources, str) else int(sources))
        h = C.c_void_p()
        _check(L.nbh_create(C.byref(self._cfg), C.byref(h)))
        self._h = h
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().nbh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def configure(self, precision=None, eps2=None, block=None, jsplit=None, ipl=None, pairing=None, sources=None):
        c = self._cfg
        if pairing is not None:
            c.pairing = PAIRINGS[pairing] if isinstance(pairing, str) else int(pairing)
        if sources is not None:
            c.sources = SOURCES[sources] if isinstance(sources, str) else int(sources)
        if ipl is not None:
            c.ipl = int(ipl)
        if precision is not None:
            c.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        if eps2 is not None:
            c.eps2 = float(eps2)
        if block is not None:
            c.block = int(block)
        if jsplit is not None:
            c.jsplit = int(jsplit)
        _check(lib().nbh_configure(self._h, C.byref(c)))

    # ---- drop-in force (force_calculation) ----
    def force(self, state, return_ms=False):
        s = _state(state)
        n = s.shape[1]
        out = np.empty((6, n))
        ms = C.c_double()
        _check(lib().nbh_force(self._h, n, *[_p(s[k]) for k in range(7)],
                               *[_p(out[k]) for k in range(6)], C.byref(ms)))
        return (out, ms.value) if return_ms else out

    def launch_shape(self, n):
        b, s, w, q, pr = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(lib().nbh_launch_shape(self._h, int(n), C.byref(b), C.byref(s), C.byref(w)))
        _check(lib().nbh_launch_ipl(self._h, int(n), C.byref(q)))
        _check(lib().nbh_launch_pairing(self._h, int(n), C.byref(pr)))
        src = C.c_int()
        _check(lib().nbh_launch_sources(self._h, int(n), C.byref(src)))
        names = {SOURCES["lds"]: "lds", SOURCES["scalar"]: "scalar"}  # 0: the symmetric kernel
        return {"block": b.value, "ipl": q.value, "jsplit": s.value, "workgroups": w.value,
                "pairing": "symmetric" if pr.value == 1 else "directed",
                "sources": names.get(src.value, "none")}

    def autotune(self, n):
        """Times launch shapes for n bodies and keeps the fastest for n's size bucket."""
        b, q, s, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        _check(lib().nbh_autotune(self._h, int(n), C.byref(b), C.byref(q), C.byref(s), C.byref(ms)))
        return {"block": b.value, "ipl": q.value, "jsplit": s.value, "trial_ms": ms.value}

    # ---- resident integrator ----
    def upload(self, state):
        s = _state(state)
        self.n = s.shape[1]
        _check(lib().nbh_upload(self._h, self.n, *[_p(s[k]) for k in range(7)]))

    def initial_dt(self, eta0=1e-4):
        d = C.c_double()
        _check(lib().nbh_initial_dt(self._h, float(eta0), C.byref(d)))
        return d.value

    def set_dt(self, dt, adaptive=False, eta=1e-2):
        _check(lib().nbh_set_dt(self._h, float(dt), int(bool(adaptive)), float(eta)))

    def get_dt(self):
        d = C.c_double()
        _check(lib().nbh_get_dt(self._h, C.byref(d)))
        return d.value

    def step(self, nsteps=1, graph=False):
        _check(lib().nbh_step(self._h, int(nsteps), int(bool(graph))))

    def download(self):
        s = np.empty((7, self.n))
        f = np.empty((6, self.n))
        _check(lib().nbh_download(self._h, *[_p(s[k]) for k in range(7)], *[_p(f[k]) for k in range(6)]))
        return s, f

    def energy(self, compat=False):
        k, u = C.c_double(), C.c_double()
        _check(lib().nbh_energy(self._h, int(bool(compat)), C.byref(k), C.byref(u)))
        return k.value, u.value

    def energy_of(self, state, compat=False):
        s = _state(state)
        k, u = C.c_double(), C.c_double()
        _check(lib().nbh_energy_host(self._h, s.shape[1], *[_p(s[i]) for i in range(7)],
                                     int(bool(compat)), C.byref(k), C.byref(u)))
        return k.value, u.value

    # ---- derivatives ----
    def force_vjp(self, state, cot):
        """d/d(state) of sum(cot * forces): state (7, n), cot (6, n) -> (7, n)."""
        s = _state(state)
        c = np.ascontiguousarray(cot, dtype=np.float64)
        if c.shape != (6, s.shape[1]):
            raise ValueError(f"cot must be (6, {s.shape[1]}), got {c.shape}")
        g = np.empty((7, s.shape[1]))
        _check(lib().nbh_force_vjp(self._h, s.shape[1], _p(s), _p(c), _p(g)))
        return g

    # ---- timing ----
    def set_profiling(self, on=True):
        _check(lib().nbh_set_profiling(self._h, int(bool(on))))

    def timing(self):
        fm, fl, sm, st = C.c_double(), C.c_int64(), C.c_double(), C.c_int64()
        _check(lib().nbh_timing(self._h, C.byref(fm), C.byref(fl), C.byref(sm), C.byref(st)))
        return {"force_ms": fm.value, "force_launches": fl.value, "step_ms": sm.value, "steps": st.value}

    def reset_timing(self):
        _check(lib().nbh_reset_timing(self._h))


# ---------------------------------------------------------------- ensemble --
ENS_BUCKETS = 4
ENS_BUCKET_EDGES = (64, 256, 1024, None)


def ensemble_sizes(nsys, nmin, nmax, seed):
    """Log-uniform integer system sizes in [nmin, nmax]:
    n_k = floor(nmin * ((nmax + 1) / nmin) ** u_k), u_k ~ U[0, 1) from
    numpy's PCG64 seeded with `seed`."""
    u = np.random.Generator(np.random.PCG64(seed)).random(nsys)
    n = np.floor(nmin * ((nmax + 1) / nmin) ** u).astype(np.int64)
    return np.clip(n, nmin, nmax).astype(np.int32)


def ensemble_state(counts, seed):
    """Concatenated Plummer spheres, system k from nbh_plummer(n_k, seed*1000003 + k),
    each in its own Henon units (total mass 1).  Returns (7, sum(counts))."""
    counts = np.asarray(counts, dtype=np.int64)
    out = np.empty((7, int(counts.sum())))
    o = 0
    L = lib()
    for k, n in enumerate(counts):
        blk = np.empty((7, int(n)))
        rc = L.nbh_plummer(int(n), int(seed) * 1000003 + k, *[_p(blk[i]) for i in range(7)])
        if rc != OK:
            raise NBodyError(rc, f"nbh_plummer(n={n}) failed")
        out[:, o:o + n] = blk
        o += n
    return out


class Ensemble:
    """nbh_ensemble: independent systems advanced by one persistent launch."""

    def __init__(self, backend, counts, eps2):
        self.counts = np.ascontiguousarray(counts, dtype=np.int32)
        self.eps2 = np.ascontiguousarray(np.broadcast_to(eps2, self.counts.shape), dtype=np.float64)
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)[:-1]]).astype(np.int64)
        self.total = int(self.counts.sum())
        self.backend = backend
        h = C.c_void_p()
        _check(lib().nbh_ens_create(backend._h, len(self.counts),
                                    self.counts.ctypes.data_as(C.POINTER(C.c_int)),
                                    _p(self.eps2), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().nbh_ens_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, state):
        s = _state(state)
        if s.shape[1] != self.total:
            raise ValueError(f"state has {s.shape[1]} bodies, ensemble {self.total}")
        _check(lib().nbh_ens_upload(self._h, *[_p(s[k]) for k in range(7)]))

    def set_dt(self, dt):
        _check(lib().nbh_ens_set_dt(self._h, float(dt)))

    def step(self, nsteps=1, graph=False):
        _check(lib().nbh_ens_step(self._h, int(nsteps), int(bool(graph))))

    def download(self):
        s = np.empty((7, self.total))
        f = np.empty((6, self.total))
        _check(lib().nbh_ens_download(self._h, *[_p(s[k]) for k in range(7)], *[_p(f[k]) for k in range(6)]))
        return s, f

    def energy(self):
        k = np.empty(len(self.counts))
        u = np.empty(len(self.counts))
        _check(lib().nbh_ens_energy(self._h, _p(k), _p(u)))
        return k, u

    def autotune(self):
        t = np.zeros(ENS_BUCKETS, dtype=np.int32)
        sp = np.zeros(ENS_BUCKETS, dtype=np.int32)
        ms = np.zeros(ENS_BUCKETS)
        _check(lib().nbh_ens_autotune(self._h, t.ctypes.data_as(C.POINTER(C.c_int)),
                                      sp.ctypes.data_as(C.POINTER(C.c_int)), _p(ms)))
        return {**self.tuning(), "bucket_ms": ms.tolist()}

    def tuning(self):
        t, sp = np.zeros(ENS_BUCKETS, dtype=np.int32), np.zeros(ENS_BUCKETS, dtype=np.int32)
        q = C.c_int()
        _check(lib().nbh_ens_tuning(self._h, t.ctypes.data_as(C.POINTER(C.c_int)),
                                    sp.ctypes.data_as(C.POINTER(C.c_int)), C.byref(q)))
        return {"tile": t.tolist(), "jsplit": sp.tolist(), "ipl": q.value}

    def set_targets_per_lane(self, ipl):
        _check(lib().nbh_ens_set_targets_per_lane(self._h, int(ipl)))

    def set_tuning(self, tiles, jsplits):
        t = np.ascontiguousarray(tiles, dtype=np.int32)
        sp = np.ascontiguousarray(jsplits, dtype=np.int32)
        _check(lib().nbh_ens_set_tuning(self._h, t.ctypes.data_as(C.POINTER(C.c_int)),
                                        sp.ctypes.data_as(C.POINTER(C.c_int))))

    def info(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(lib().nbh_ens_info(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return {"items": a.value, "workgroups": b.value, "max_tile": c.value}

    def system(self, arr, k):
        """Slice of a concatenated (c, total) array belonging to system k."""
        o = self.offsets[k]
        return arr[:, o:o + self.counts[k]]


# ----------------------------------------------------------------- adjoint --
class Adjoint:
    """nbh_adjoint: checkpointed reverse sweep of a fixed-dt Hermite run."""

    def __init__(self, backend, n, every):
        self.n = int(n)
        self.backend = backend
        h = C.c_void_p()
        _check(lib().nbh_adj_create(backend._h, self.n, int(every), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().nbh_adj_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def forward(self, state0, dt, nsteps):
        s = _state(state0)
        out = np.empty((7, self.n))
        _check(lib().nbh_adj_forward(self._h, _p(s), float(dt), int(nsteps), _p(out)))
        return out

    def backward(self, cot_final):
        c = np.ascontiguousarray(cot_final, dtype=np.float64)
        g = np.empty((7, self.n))
        _check(lib().nbh_adj_backward(self._h, _p(c), _p(g)))
        return g
