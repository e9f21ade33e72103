"""The C ABI library loads, exports what include/nbody_hip.h declares, and
its host-only parts (initial conditions, argument checks) behave.  CPU only:
no call here launches device work."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, golden

HEADER = os.path.join(ROOT, "include", "nbody_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nbh_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("nbh_force", "nbh_create", "nbh_destroy", "nbh_step", "nbh_upload",
                 "nbh_energy", "nbh_plummer", "nbh_reference_ic", "nbh_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol(nb):
    L = nb.lib()
    missing = [s for s in declared() if not hasattr(L, s)]
    assert not missing, missing


def test_version(nb):
    assert nb.lib().nbh_version() == 104


def test_reference_ic_matches_reference_fixture(nb):
    g = golden("ref_ic_n1024.npz")
    assert np.array_equal(nb.reference_ic(1024), g["state"])


def test_reference_ic_matches_oracle_other_sizes(nb, oracle):
    for n in (1, 2, 17, 300):
        # n = 1: U = 0, Q = -inf, the rescale makes NaN velocities, as in the reference
        assert np.array_equal(nb.reference_ic(n), oracle.reference_ic(n), equal_nan=True)


def test_plummer_is_deterministic_and_seeded(nb):
    a = nb.plummer(1000, 5)
    assert np.array_equal(a, nb.plummer(1000, 5))
    assert not np.array_equal(a, nb.plummer(1000, 6))
    assert np.all(a[3] == 1.0 / 1000)


def test_plummer_henon_units(nb, oracle):
    s = nb.plummer(20000, 1)
    m = s[3]
    for c in (0, 1, 2, 4, 5, 6):  # centre-of-mass frame
        assert abs(np.sum(m * s[c])) < 1e-12
    T = oracle.kinetic(s)
    U = oracle.potential(s, 0.0, nthreads=8)
    assert abs((T + U) - (-0.25)) < 0.01   # E = -1/4 up to truncation and N noise
    assert abs(-2 * T / U - 1.0) < 0.03    # virial equilibrium
    r = np.sqrt(s[0] ** 2 + s[1] ** 2 + s[2] ** 2)
    a = 3 * np.pi / 16
    # Plummer half-mass radius a / sqrt(2^(2/3) - 1)
    assert abs(np.median(r) / (a / np.sqrt(2 ** (2 / 3) - 1)) - 1) < 0.03
    # truncation at mass fraction 0.999: r_max = a / sqrt(0.999^(-2/3) - 1)
    assert r.max() <= a / np.sqrt(0.999 ** (-2 / 3) - 1) * (1 + 1e-12)


def test_ic_reject_bad_arguments(nb):
    L = nb.lib()
    buf = np.zeros(4)
    p = buf.ctypes.data_as(C.POINTER(C.c_double))
    assert L.nbh_plummer(0, 1, p, p, p, p, p, p, p) == nb.ERR_ARG
    assert L.nbh_plummer(4, 1, p, p, p, None, p, p, p) == nb.ERR_ARG
    assert L.nbh_reference_ic(-1, 1.0, 1.0, 0.01, 1.0, p, p, p, p, p, p, p) == nb.ERR_ARG


def test_null_context_is_an_argument_error(nb):
    L = nb.lib()
    assert L.nbh_force(None, 4, *([None] * 13), None) == nb.ERR_ARG
    assert b"context" in L.nbh_last_error()
    assert L.nbh_step(None, 1, 0) == nb.ERR_ARG
    assert L.nbh_destroy(None) == nb.OK


def test_bad_config_rejected_before_device_use(nb):
    L = nb.lib()
    h = C.c_void_p()
    for cfg in (nb.Config(0, 7, 1e-7, 0, 0), nb.Config(0, 0, -1.0, 0, 0),
                nb.Config(0, 0, 1e-7, 96, 0), nb.Config(0, 0, 1e-7, 0, -2),
                nb.Config(0, 1, 1e-7, 0, 0, 0, 2),     # unknown pairing
                nb.Config(0, 1, 1e-7, 0, 0, 0, 0, 3)):  # unknown source path
        assert L.nbh_create(C.byref(cfg), C.byref(h)) == nb.ERR_ARG
    assert b"sources must be" in L.nbh_last_error()
    assert L.nbh_create(None, C.byref(h)) == nb.ERR_ARG


def test_no_device_fails_loudly(nb):
    """Without a GPU the product refuses to run; it never computes on the CPU."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(nb.NBodyError) as e:
        nb.Backend("fp64", 1e-7)
    assert e.value.code == nb.ERR_HIP
