"""Body-state files (nbh_state_save / nbh_state_info / nbh_state_load): host
code, CPU only."""
import os
import struct

import numpy as np
import pytest


def test_round_trip_is_bitwise(nb, tmp_path):
    s = nb.plummer(777, 3)
    p = tmp_path / "s.nbh"
    nb.state_save(p, s, time=0.25, step=128)
    assert nb.state_info(p) == {"n": 777, "time": 0.25, "step": 128}
    assert np.array_equal(nb.state_load(p), s)
    assert os.path.getsize(p) == 64 + 7 * 8 * 777


def test_header_layout(nb, tmp_path):
    p = tmp_path / "h.nbh"
    nb.state_save(p, nb.plummer(5, 1), time=1.5, step=7)
    raw = open(p, "rb").read()
    magic, version, reserved, n, t, step = struct.unpack("<8sIIqdq", raw[:40])
    assert (magic, version, reserved, n, t, step) == (b"NBHSTATE", 1, 0, 5, 1.5, 7)
    assert raw[40:64] == bytes(24)


@pytest.mark.parametrize("damage", ["magic", "truncate", "version", "empty"])
def test_bad_files_are_argument_errors(nb, tmp_path, damage):
    p = tmp_path / "bad.nbh"
    nb.state_save(p, nb.plummer(10, 2))
    raw = bytearray(open(p, "rb").read())
    if damage == "magic":
        raw[0:8] = b"NOTSTATE"
    elif damage == "truncate":
        raw = raw[:-8]
    elif damage == "version":
        raw[8:12] = struct.pack("<I", 9)
    else:
        raw = raw[:10]
    open(p, "wb").write(bytes(raw))
    with pytest.raises(nb.NBodyError) as e:
        nb.state_load(p)
    assert e.value.code == nb.ERR_ARG


def test_missing_file(nb, tmp_path):
    with pytest.raises(nb.NBodyError) as e:
        nb.state_info(tmp_path / "nope.nbh")
    assert b"cannot open" in nb.lib().nbh_last_error()
