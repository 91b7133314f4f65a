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


# ---- the SCEPTIC reinjection's distributions (pot.c:496-652) -------------------
#
# Without drift and at chi = 0, the speed at infinity u (in units of
# v_n = sqrt(2 kT/m)) follows Hutchinson's (2003) eq. 19 at U = 0,
#     F(u) = 1 - (u^2 + 1 - chi) exp(-u^2) / (1 - chi)
# (prng.c:138-154, pot.c:546-562); the impact parameter b has b^2 uniform
# on [0, 1 - chi/u^2] (pot.c:582); energy and angular momentum conservation
# give the boundary velocity (pot.c:617-633), so |v|^2/v_n^2 = u^2 - chi
# and (v.r_hat)^2 / |v|^2 is uniform on [0, 1]; and with no preferred
# direction the position is uniform on the sphere.  At chi != 0 a sample
# whose orbit meets a potential barrier (alpha NaN) is drawn again
# (pot.c:584-599), which reweights u and b by a factor with no closed form;
# there the position must still be isotropic and every particle must have
# u^2 >= max(chi, 0).  A dropped term in the energy balance, a wrong sign
# of chi or a transposed rotation breaks one of these.

def no_flow_cdf(u, chi):
    return 1.0 - (u * u + 1.0 - chi) * np.exp(-u * u) / (1.0 - chi)


def conditional_speed_cdf(u, chi):
    """F conditioned on u^2 >= chi"""
    lo = no_flow_cdf(np.sqrt(max(chi, 0.0)), chi) if chi > 0 else 0.0
    return (no_flow_cdf(u, chi) - lo) / (1.0 - lo)


def reinjection_invariants(parts, R, v_n, chi_s):
    """(u-CDF values, (v.r_hat)^2/|v|^2, cos polar angle, azimuth/2pi) of
    reinjected particles at positions (x, y, z) on the sphere of radius R
    about (R, R, R)"""
    d = np.stack([parts["x"] - R, parts["y"] - R, parts["z"] - R], 1)
    r = np.linalg.norm(d, axis=1)
    v = np.stack([parts["vx"], parts["vy"], parts["vz"]], 1)
    v2 = np.sum(v * v, axis=1)
    u = np.sqrt(v2 / (v_n * v_n) + chi_s)
    cos_vr2 = (np.sum(v * d, axis=1) / r) ** 2 / v2
    return (conditional_speed_cdf(u, chi_s), cos_vr2, d[:, 2] / r,
            (np.arctan2(d[:, 1], d[:, 0]) / (2 * np.pi)) % 1.0, np.sum(v * d, axis=1))


@pytest.mark.parametrize("chi_b", [0.0, 0.3, -0.6])
@pytest.mark.parametrize("species", [O.ION, O.ELECTRON], ids=["ion", "electron"])
def test_oracle_reinjection_distributions(chi_b, species):
    from scipy import stats
    case = CASES[0]  # sphere, Te = Ti, no drift
    sim = oracle_sim(case)
    L = sim.L
    rng = np.random.default_rng(17 + species)
    n = 6000
    out = np.zeros(n, O.PARTICLE_DT)
    p = np.zeros(1, O.PARTICLE_DT)
    for k in range(n):
        p[0]["species"] = species
        u = np.ascontiguousarray(rng.random(1024))
        used = L.or_sceptic_from_uniforms(sim.h, p.ctypes.data, chi_b, O.dptr(u), 1024)
        assert used > 0
        out[k] = p[0]
    R = float(case["args"]["-r"]) * float(case["args"]["-m"])
    T = float(case["args"]["-E"] if species == O.ELECTRON else case["args"]["-I"])
    m = 9.109383e-31 * (1.0 if species == O.ELECTRON else 1836.15)
    v_n = np.sqrt(2 * 1.38065e-23 * T / m)
    # electrons see chi_b scaled by -Ti/Te (pot.c:536-539)
    chi_s = chi_b if species == O.ION else -chi_b * float(case["args"]["-I"]) / float(
        case["args"]["-E"])
    Fu, cos_vr2, cz, az, vdotr = reinjection_invariants(out, R, v_n, chi_s)
    r = np.sqrt((out["x"] - R) ** 2 + (out["y"] - R) ** 2 + (out["z"] - R) ** 2)
    assert np.allclose(r, R, rtol=1e-12)
    assert np.all(vdotr < 0)  # every particle enters the domain
    v2 = out["vx"] ** 2 + out["vy"] ** 2 + out["vz"] ** 2
    assert np.all(v2 / v_n ** 2 + chi_s >= max(chi_s, 0.0) * (1 - 1e-12))
    checks = [("cos polar", (cz + 1) / 2), ("azimuth", az)]
    if chi_b == 0.0:
        checks += [("speed", Fu), ("radial fraction", cos_vr2)]
    for name, x in checks:
        pval = stats.kstest(x, "uniform").pvalue
        assert pval > 1e-3, (name, pval)
