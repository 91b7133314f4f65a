"""The product's host-side initial load (pot_amd/host/pot_init.c) against
the oracle's init_particles, which the reference-run fixtures pin: same seed,
same particles bit for bit (pot.c:887-940, prng.c:4-47)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from gpu_common import CASES, big_case, gpu_config, oracle_sim
from pot_amd import potg as G

HOSTLIB = os.path.join(os.path.dirname(G.LIB_PATH), "libpothost.so")


def host_init(case, seed_words):
    L = C.CDLL(HOSTLIB)
    cfg = gpu_config(case)
    N = cfg.N
    parts = np.zeros(N, G.PARTICLE_DT)
    st = (C.c_uint32 * 21)()
    rc = L.pot_init_particles(C.byref(cfg), (C.c_uint32 * 16)(*seed_words),
                              parts.ctypes.data_as(C.c_void_p), st)
    assert rc == 0
    return parts, np.array(st, np.uint32)


@pytest.mark.parametrize("case", CASES + [big_case(20001, geometry=1, A=2.0)],
                         ids=[c["name"] for c in CASES] + ["big_prolate"])
def test_host_init_matches_oracle(case):
    from make_golden import splitmix64_words
    w = splitmix64_words(case["seed"])
    parts, st = host_init(case, w)
    ref = oracle_sim(case).particles()
    for f in ("x", "y", "z", "vx", "vy", "vz", "species"):
        assert np.array_equal(parts[f], ref[f]), f
    sim = oracle_sim(case)
    pr = O.Prng()
    sim.L.or_get_prng(sim.h, C.byref(pr))
    assert list(st[4:]) == [pr.index] + list(pr.state)
    assert list(st[:4]) == [pr.a, pr.b, pr.c, pr.d]
