"""The edge potential chi_b, the potential at points, and reinjection at
chi_b != 0: the HIP path against the oracle on the same snapshots.

- chi_b is measured every 20 steps, before any particle moves
  (measure_chi_b, pot.c:2240-2280, called at pot.c:2409-2418): the
  potential at one point of the domain's edge (azimuth 0, on the equator
  through the grain) over kT_i/e.  A context uploaded with 19 steps done
  measures it in its next step; that step's reinjections then use it.
  Spheroid programs add the grain's potential at the point's absolute
  coordinates, not relative to the grain (SPH:1929-1934) -- a quirk the
  oracle keeps and the GPU must too.
- potg_phi_at is phi_at (pot.c:1818-1901; SPH:1922-2035), the float64 walk
  with every leaf counted.
- sceptically_reinject (pot.c:496-652) feeds chi_b into the speed
  inverse-CDFs (prng.c:138-214) and into integrate_for_alpha
  (pot.c:433-492); uploads with chi_b = -0.6 and 0.3 reach those branches.
  (An electron sees -chi_b Ti/Te, pot.c:536-539; chi = 1 exactly makes the
  no-flow inverse CDF divide by zero, prng.c:139-152, in the reference as
  here, so the cases stay clear of it.)
- the reinjected particles' distributions: speed, entry angle and position
  against the closed forms the oracle's sampler is checked against.

The GPU evaluates the same float64 expressions in the same order; the
compiler may fuse a multiply and an add, so sums agree to 1e-12 relative.
"""
import numpy as np
import pytest

from gpu_common import (CASES, assert_particles_match, assert_step_outputs_equal, big_case,
                        domain_radius, gpu_ctx, oracle_sim, step_both)

from pot_amd import potg as G

pytestmark = pytest.mark.gpu

BIG = [big_case(20001, seed=41), big_case(20001, geometry=1, A=2.0, seed=42),
       big_case(20001, geometry=2, A=0.5, seed=43),
       dict(big_case(20001, seed=44), name="big20001_dipoles", dipoles=1)]
ALL = CASES + BIG
IDS = [c["name"] for c in ALL]


@pytest.mark.parametrize("case", ALL, ids=IDS)
def test_chi_b_at_step_20_matches_oracle(case):
    """Upload after 19 steps: step 20 measures chi_b from the step's tree
    and reinjects with it.  chi_b to 1e-12, then the whole step."""
    sim, ctx, oo, og, p0 = step_both(case, steps=1, t0=19)
    want = sim.state()["chi_b"]
    _, gr = ctx.download(0)
    assert want != 0.0
    assert abs(gr["chi_b"] - want) <= 1e-12 * abs(want), (gr["chi_b"], want)
    assert_step_outputs_equal(oo[0], og[0])
    pg, _ = ctx.download(0)
    assert_particles_match(case, sim.particles(), pg, p0)
    assert gr["q"] == sim.state()["q"]


@pytest.mark.parametrize("case", [c for c in ALL if c["name"] in
                                  ("sphere", "prolate", "oblate", "sphere_dipoles")] + BIG,
                         ids=lambda c: c["name"])
def test_chi_b_over_two_measurements(case):
    """Steps 20 and 40 both measure chi_b: after 21 steps from t = 19 the
    second value, measured from the GPU's own state, still matches the
    oracle's (the integrators agree to 1e-12 per step)."""
    sim, ctx, oo, og, p0 = step_both(case, steps=21, t0=19)
    want = sim.state()["chi_b"]
    _, gr = ctx.download(0)
    assert abs(gr["chi_b"] - want) <= 1e-9 * abs(want), (gr["chi_b"], want)
    for a, b in zip(oo, og):
        assert_step_outputs_equal(a, b)


def _points(case, parts, n, seed):
    """sample points: the chi_b point, random points inside the domain's
    sphere, points next to particles, and the grain's own position"""
    rng = np.random.default_rng(seed)
    R = domain_radius(case)
    g = np.array([parts["x"][0], parts["y"][0], parts["z"][0]])
    pts = [g + np.array([R, 0.0, 0.0]), g]
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    r = R * rng.uniform(0.0, 1.0, size=n) ** (1.0 / 3.0)
    pts += list(R + u * r[:, None])
    j = rng.integers(1, len(parts), size=n)
    near = np.stack([parts["x"][j], parts["y"][j], parts["z"][j]], 1)
    pts += list(near + rng.normal(scale=1e-3 * R, size=(n, 3)))
    return np.array(pts)


@pytest.mark.parametrize("case", ALL, ids=IDS)
def test_phi_at_matches_oracle(case):
    sim = oracle_sim(case)
    parts = sim.particles()
    sim.build_tree()
    pts = _points(case, parts, 400, 3)
    want = np.array([sim.phi_at(*p) for p in pts])
    for prec in (G.POTG_FP64, G.POTG_MIXED):
        ctx = gpu_ctx(case, parts, precision=prec)
        ctx.build_tree()
        got = ctx.phi_at(0, pts)
        scale = np.abs(want).max()
        err = np.abs(got - want)
        assert np.all(err <= 1e-12 * np.abs(want) + 1e-14 * scale), (prec, err.max() / scale)


def test_phi_at_spheroid_uses_absolute_coordinates():
    """SPH:1929-1934 evaluates the grain's potential at (x, y, z) itself,
    not at its offset from the grain; with the grain at (R, R, R) that is a
    point far away, so the grain term at the edge is small but not zero.
    The GPU must keep the quirk: moving the grain charge changes phi_at by
    exactly what the oracle's absolute-coordinate term changes by."""
    case = next(c for c in CASES if c["name"] == "prolate")
    sim = oracle_sim(case)
    parts = sim.particles()
    R = domain_radius(case)
    pt = np.array([[2.0 * R, R, R]])
    vals = []
    for q in (0, -500):
        st = sim.state()
        sim.set_state(q, st["m"], 0, 0.0, 0)
        sim.build_tree()
        ctx = gpu_ctx(case, parts, q=q)
        ctx.build_tree()
        vals.append((ctx.phi_at(0, pt)[0], sim.phi_at(*pt[0])))
    d_gpu = vals[1][0] - vals[0][0]
    d_ora = vals[1][1] - vals[0][1]
    assert d_ora != 0.0
    assert abs(d_gpu - d_ora) <= 1e-10 * abs(d_ora), (d_gpu, d_ora)


@pytest.mark.parametrize("chi_b", [-0.6, 0.3], ids=["chi_m0.6", "chi_0.3"])
@pytest.mark.parametrize("case", [c for c in CASES if c["name"] in
                                  ("sphere", "sphere_bx_drift", "prolate", "oblate",
                                   "sphere_dipoles")] + BIG[:2],
                         ids=lambda c: c["name"])
def test_reinjection_at_nonzero_chi_b(case, chi_b):
    """One step at a fixed chi_b != 0 (not a measuring step): every
    reinjection samples u with u^2 >= chi_b (scaled by -Ti/Te for electrons,
    pot.c:546-560), integrates alpha with chi_b (pot.c:433-492) and sets the
    velocity from energy conservation (pot.c:617-633).  Ten steps of the
    small cases (t = 1..10, before the first measuring step), so that every
    case reinjects some particles: the small prolate case first reinjects at
    its fifth step.  Four of the 20001-particle ones: over more steps one
    of their many reinjections can come out grazing the boundary, where the
    last bit of a float64 sine decides whether the next push leaves it
    inside, and GPU and host libraries round such bits differently."""
    steps = 4 if len(case["args"]) and int(case["args"]["-N"]) > 10000 else 10
    sim, ctx, oo, og, p0 = step_both(case, steps=steps, t0=0, chi_b=chi_b, seed=91)
    n_re = 0
    for a, b in zip(oo, og):
        assert_step_outputs_equal(a, b)
        n_re += a.n_reinj_abs + a.n_reinj_out
    assert n_re > 0
    pg, gr = ctx.download(0)
    assert gr["chi_b"] == chi_b
    assert_particles_match(case, sim.particles(), pg, p0, x_tol=1e-11, v_tol=1e-9)
    assert gr["q"] == sim.state()["q"]


@pytest.mark.parametrize("chi_b", [0.0, 0.3], ids=["chi_0", "chi_0.3"])
def test_reinjected_electrons_follow_the_closed_forms(chi_b):
    """Electrons reinjected by the GPU over 18 steps (t = 2..19, chi_b held
    at the uploaded value): each lies on the domain's sphere moving inwards,
    positions are isotropic, and at chi_b = 0 the speed and the entry angle
    follow the closed forms of the oracle's sampler test (Hutchinson 2003
    eq. 19 at U = 0; (v.r_hat)^2/|v|^2 uniform)."""
    from scipy import stats
    from test_oracle_properties import reinjection_invariants
    case = big_case(200001, seed=45)
    parts = oracle_sim(case).particles()
    ctx = gpu_ctx(case, parts, seed=13, t=1, chi_b=chi_b)
    R = domain_radius(case)
    el = parts["species"] == G.ELECTRON
    got = []
    prev, _ = ctx.download(0)
    for _ in range(18):
        ctx.step()
        p, _ = ctx.download(0)
        r = np.sqrt((p["x"] - R) ** 2 + (p["y"] - R) ** 2 + (p["z"] - R) ** 2)
        moved = np.sqrt((p["x"] - prev["x"]) ** 2 + (p["y"] - prev["y"]) ** 2 +
                        (p["z"] - prev["z"]) ** 2) > 0.05 * R
        on = el & moved & (np.abs(r - R) <= 1e-9 * R)
        got.append(p[on].copy())
        prev = p
    got = np.concatenate(got)
    assert len(got) > 1000, len(got)
    Te = float(case["args"]["-E"])
    v_n = np.sqrt(2 * 1.38065e-23 * Te / 9.109383e-31)
    chi_s = -chi_b * float(case["args"]["-I"]) / Te
    Fu, cos_vr2, cz, az, vdotr = reinjection_invariants(got, R, v_n, chi_s)
    assert np.all(vdotr < 0)
    checks = [("cos polar", (cz + 1) / 2), ("azimuth", az)]
    if chi_b == 0.0:
        checks += [("speed", Fu), ("radial fraction", cos_vr2)]
    for name, x in checks:
        pval = stats.kstest(x, "uniform").pvalue
        assert pval > 1e-3, (name, pval, len(got))
