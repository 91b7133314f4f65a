"""The product's state-file reader and writer (pot_amd/host/pot_state.c)
against state files the reference binary wrote (tests/golden/*/state_*.bin,
format C, pot.c:3816-4019): reading one and writing it back reproduces it
byte for byte, and the particles and PRNG record read equal the file's."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
from gpu_common import CASES, GOLD
from pot_amd import potg as G

HOSTLIB = os.path.join(os.path.dirname(G.LIB_PATH), "libpothost.so")


class Prng(C.Structure):
    _fields_ = [("a", C.c_uint32), ("b", C.c_uint32), ("c", C.c_uint32), ("d", C.c_uint32),
                ("index", C.c_uint32), ("state", C.c_uint32 * 16)]


class SimState(C.Structure):
    _fields_ = [("grain_q", C.c_int32), ("tot_absor_e", C.c_uint32), ("grain_m", C.c_longdouble),
                ("t_completed", C.c_uint64), ("epe", C.c_longdouble * 3), ("chi_b", C.c_double),
                ("prng", Prng)]


def _lib():
    L = C.CDLL(HOSTLIB)
    L.pot_read_state.argtypes = [C.c_char_p, C.c_uint64, C.c_void_p, C.POINTER(SimState)]
    L.pot_write_state.argtypes = [C.c_char_p, C.c_uint64, C.c_void_p, C.POINTER(SimState),
                                  C.c_int]
    return L


@pytest.mark.parametrize("which", ["state_1.bin", "state_K.bin"])
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_state_roundtrip_is_byte_identical(case, which, tmp_path):
    L = _lib()
    N = int(case["args"]["-N"])
    src = os.path.join(GOLD, case["name"], which)
    parts = np.zeros(N, G.PARTICLE_DT)
    st = SimState()
    raw = open(src, "rb").read()
    if raw[raw.index(b"\n") + 1] in b" \t\n\v\f\r":
        # the reference's reader ends its header format with "\n"
        # (pot.c:3923), which skips whitespace, so a binary PRNG record that
        # starts with a whitespace byte is misread and the particle read
        # comes up short: the reference cannot resume from such a file, and
        # neither can this reader
        assert L.pot_read_state(src.encode(), N, parts.ctypes.data, C.byref(st)) != 0
        return
    assert L.pot_read_state(src.encode(), N, parts.ctypes.data, C.byref(st)) == 0
    hdr, prng, ref_parts = O.read_state_file(src, N)
    for f in ("x", "y", "z", "vx", "vy", "vz", "species"):
        assert np.array_equal(parts[f], ref_parts[f]), f
    assert [st.prng.index] + list(st.prng.state) == list(prng[4:])
    steps = 1 if which == "state_1.bin" else case["steps"]
    assert st.t_completed == steps
    out = tmp_path / "s.bin"
    assert L.pot_write_state(str(out).encode(), N, parts.ctypes.data, C.byref(st), 1) == 0
    assert out.read_bytes() == open(src, "rb").read()


def test_state_rejects_other_worker_counts(tmp_path):
    L = _lib()
    src = os.path.join(GOLD, CASES[0]["name"], "state_1.bin")
    raw = open(src, "rb").read()
    assert raw.startswith(b"C 1 ")
    bad = tmp_path / "two.bin"
    bad.write_bytes(b"C 2 " + raw[4:])
    N = int(CASES[0]["args"]["-N"])
    parts = np.zeros(N, G.PARTICLE_DT)
    st = SimState()
    assert L.pot_read_state(str(bad).encode(), N, parts.ctypes.data, C.byref(st)) != 0
    junk = tmp_path / "junk.bin"
    junk.write_bytes(b"xyz")
    assert L.pot_read_state(str(junk).encode(), N, parts.ctypes.data, C.byref(st)) != 0
