"""The product's macroscopic-record writer (pot_amd/host/pot_macro.c)
against the reference's own pot-macroscopic.dat files.

tests/golden/*/macro.dat were written by the reference binary on steps
1..K (every step up to 200 is a record step, pot.c:4364-4369).  The oracle,
run in reference mode, reproduces the reference's states bit for bit
(test_oracle_golden.py); each record is formatted here from the oracle's
state after the step -- particles, grain charge and mass, electrons
absorbed, and the potential energies the oracle's compute_epe pass summed
(phi_particle_feels, pot.c:1934-2018, 2771-2863) -- and must equal the
reference's line character for character.  This pins the writer's 35
columns and the oracle's potential energies (PE, PI, PG columns) at the
printed precision.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
from gpu_common import CASES, GOLD
from make_golden import splitmix64_words
from pot_amd import potg as G

HOSTLIB = os.path.join(os.path.dirname(G.LIB_PATH), "libpothost.so")


class MacroIn(C.Structure):
    _fields_ = [("parts", C.c_void_p), ("N", C.c_uint64), ("r", C.c_uint32), ("mpp", C.c_double),
                ("dt", C.c_longdouble), ("mime", C.c_double), ("grain_m", C.c_longdouble),
                ("grain_q", C.c_int32), ("tot_absor_e", C.c_uint32),
                ("epe", C.c_longdouble * 3), ("time_step", C.c_uint64)]


def _lib():
    L = C.CDLL(HOSTLIB)
    L.pot_macro_format.argtypes = [C.POINTER(MacroIn), C.c_char_p, C.c_size_t]
    L.pot_macro_format.restype = C.c_int
    L.pot_macro_header.restype = C.c_char_p
    return L


def format_record(L, sim, case):
    parts = np.ascontiguousarray(sim.particles())
    q = C.c_int32()
    m = C.c_longdouble()
    t = C.c_uint64()
    tot = C.c_uint32()
    epe = (C.c_longdouble * 3)()
    sim.L.or_get_state(sim.h, C.byref(q), C.byref(m), C.byref(t), None, C.byref(tot), epe)
    a = case["args"]
    mi = MacroIn()
    mi.parts = parts.ctypes.data
    mi.N = len(parts)
    mi.r = int(a["-r"])
    mi.mpp = float(a["-m"])
    mi.dt = sim.cfg.dt
    mi.mime = 1836.153 if case["geometry"] else 1836.15
    mi.grain_m = m.value
    mi.grain_q = q.value
    mi.tot_absor_e = tot.value
    for k in range(3):
        mi.epe[k] = epe[k]
    mi.time_step = t.value
    buf = C.create_string_buffer(4096)
    n = L.pot_macro_format(C.byref(mi), buf, len(buf))
    assert n > 0
    return buf.value.decode()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_macro_records_match_reference(case):
    L = _lib()
    lines = open(os.path.join(GOLD, case["name"], "macro.dat")).read().splitlines(True)
    assert lines[0] == L.pot_macro_header().decode()
    cfg = O.make_config(case["args"], geometry=case["geometry"], theta=case["theta"],
                        B=tuple(case["B"]), dipoles=case.get("dipoles", 0),
                        integrator=case.get("integrator", 0), bounce=case.get("bounce"))
    sim = O.Sim(cfg)
    sim.init_particles(splitmix64_words(case["seed"]))
    steps = min(case["steps"], len(lines) - 1)
    for t in range(1, steps + 1):
        sim.run_step()
        got = format_record(L, sim, case)
        assert got == lines[t], "step %d:\n got %s want %s" % (t, got, lines[t])


def test_macro_needs_six_particles():
    L = _lib()
    parts = np.zeros(5, G.PARTICLE_DT)
    mi = MacroIn()
    mi.parts = parts.ctypes.data
    mi.N = 5
    buf = C.create_string_buffer(64)
    assert L.pot_macro_format(C.byref(mi), buf, len(buf)) == 0
