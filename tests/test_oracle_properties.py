"""Checks of the oracle that do not come from reference runs: closed forms,
published known answers and limits, chosen so that a dropped term, a wrong
sign or a wrong index in the oracle fails one of them.

- Philox4x32-10 against the published known-answer vectors (Salmon et al.,
  SC'11; Random123 kat_vectors), the generator of the GPU-semantics
  reinjection stream.
- Walk at theta = 0: no cell passes the MAC (pot.c:1727), so the walk is the
  softened direct sum over the same sources (pot.c:2213-2233), up to
  summation order -- with monopoles and with DIPOLES.
- integrate_for_alpha at chi_b = 0 (pot.c:433-492) is the integral of
  1/sqrt(1/s^2 - xi^2) over [0, 1], i.e. asin(s).
- u_no_flow_icdf (prng.c:137-154) inverts 1 - (u^2 + 1 - chi) e^{-u^2}/(1 - chi);
  its chi = 0 median, found here by bisection, is 1.29551 (the reference's
  Newton starting guess, 1.29951 at prng.c:142, is close to it, not equal).
- u_icdf (prng.c:168-214) tends to the no-flow inverse as the drift U -> 0
  (a different formula, so the two check each other).
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O
from gpu_common import CASES, oracle_sim


def philox(ctr, key):
    L = O.lib()
    out = (C.c_uint32 * 4)()
    L.or_philox4x32_10((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
    return list(out)


def test_philox4x32_10_known_answers():
    assert philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                           0x6D5451FD]
    assert philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                  [0xA4093822, 0x299F31D0]) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


@pytest.mark.parametrize("dipoles", [0, 1])
@pytest.mark.parametrize("case", [CASES[0], CASES[4]], ids=["sphere", "prolate"])
def test_walk_at_theta_zero_is_the_direct_sum(case, dipoles):
    cfg = O.make_config(case["args"], geometry=case["geometry"], theta=0.0, inplace=0)
    cfg.dipoles = dipoles
    sim = O.Sim(cfg)
    from make_golden import splitmix64_words
    sim.init_particles(splitmix64_words(case["seed"]))
    sim.build_tree()
    p = sim.particles()
    for i in range(1, len(p), 37):
        x, y, z = p["x"][i], p["y"][i], p["z"][i]
        Ew = sim.field_at(i, x, y, z, 1)
        Ed = sim.direct_sum(i, x, y, z)
        assert np.linalg.norm(Ew - Ed) <= 1e-12 * np.linalg.norm(Ed), i


def test_dipoles_reduce_the_truncation_error():
    """At theta = 0.5 the dipole correction (pot.c:1732-1750) must bring the
    walk closer to the direct sum on average; a wrong sign in the dipole
    term makes it worse."""
    from make_golden import splitmix64_words
    case = CASES[0]
    errs = []
    for dip in (0, 1):
        cfg = O.make_config(case["args"], geometry=0, theta=0.5, inplace=0)
        cfg.dipoles = dip
        sim = O.Sim(cfg)
        sim.init_particles(splitmix64_words(case["seed"]))
        sim.build_tree()
        p = sim.particles()
        e = []
        for i in range(1, len(p), 3):
            x, y, z = p["x"][i], p["y"][i], p["z"][i]
            Ed = sim.direct_sum(i, x, y, z)
            e.append(np.linalg.norm(sim.field_at(i, x, y, z, 1) - Ed) / np.linalg.norm(Ed))
        errs.append(np.median(e))
    assert errs[1] < 0.5 * errs[0], errs


@pytest.mark.parametrize("s", [1e-3, 0.1, 0.3, 0.5, 0.7, 0.9])
def test_integrate_for_alpha_without_potential_is_asin(s):
    L = O.lib()
    a = L.or_integrate_for_alpha(s, 0.0, 1.0, 3.0, 0.01)
    assert abs(a - math.asin(s)) < 2e-3 * math.asin(s) + 1e-9


def test_u_no_flow_icdf_inverts_its_cdf():
    L = O.lib()
    lo, hi = 0.5, 3.0
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if 1.0 - (mid * mid + 1.0) * math.exp(-mid * mid) < 0.5:
            lo = mid
        else:
            hi = mid
    assert abs(L.or_u_no_flow_icdf(0.0, 0.5) - lo) < 1e-7
    assert abs(lo - 1.29551) < 1e-5
    for chi in (0.0, -0.5, -2.0):
        for q in (0.01, 0.2, 0.5, 0.8, 0.99):
            u = L.or_u_no_flow_icdf(chi, q)
            cdf = 1.0 - (u * u + 1.0 - chi) * math.exp(-u * u) / (1.0 - chi)
            assert abs(cdf - q) < 1e-7, (chi, q)


def test_u_icdf_tends_to_no_flow_as_drift_vanishes():
    L = O.lib()
    for q in (0.05, 0.3, 0.5, 0.7, 0.95):
        a = L.or_u_icdf(1e-4, 0.0, q)
        b = L.or_u_no_flow_icdf(0.0, q)
        assert abs(a - b) < 1e-3, (q, a, b)
    prev = 0.0
    for q in (0.1, 0.3, 0.5, 0.7, 0.9):  # monotone in q at finite drift
        u = L.or_u_icdf(1.0, 0.0, q)
        assert u > prev
        prev = u


def test_frozen_and_live_sources_agree_before_anything_moves():
    """GPU semantics read the build snapshot; reference semantics read the
    live array (pot.c:1732-1790).  Before a step moves anything the two are
    the same array, so the walks must agree exactly."""
    case = CASES[1]
    a = oracle_sim(case, inplace=0)
    b = oracle_sim(case, inplace=1)
    a.build_tree()
    b.build_tree()
    p = a.particles()
    for i in range(1, len(p), 29):
        x, y, z = p["x"][i], p["y"][i], p["z"][i]
        assert np.array_equal(a.field_at(i, x, y, z, 0), b.field_at(i, x, y, z, 0))
