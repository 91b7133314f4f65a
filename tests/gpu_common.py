"""Shared set-up of the GPU parity tests: oracle-initialised states and the
matching device contexts."""
import json
import os

import numpy as np

import oracle_lib as O
from make_golden import plasma_args, splitmix64_words

from pot_amd import potg as G

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))


def big_case(N, geometry=0, A=1.0, te=11600.0, ti=11600.0, B=(0.0, 0.0, 0.0), seed=11,
             r_over_ld=8.0, a_over_ld=0.5, drift_mach=0.0):
    args = plasma_args(N, r_over_ld, te, ti, a_over_ld, 0.01, 0.01)
    if geometry:
        args["-A"] = repr(A)
    if drift_mach:
        args["-x"] = "%.6g" % (drift_mach * np.sqrt(1.38065e-23 * te / (1836.15 * 9.109383e-31)))
    return {"name": "big%d_g%d" % (N, geometry), "geometry": geometry, "theta": 0.5,
            "B": list(B), "args": args, "seed": seed, "steps": 1}


def oracle_sim(case, inplace=0, rng_kind=O.RNG_WELL512, philox_seed=0):
    cfg = O.make_config(case["args"], geometry=case["geometry"], theta=case["theta"],
                        B=tuple(case["B"]), inplace=inplace, rng_kind=rng_kind,
                        philox_seed=philox_seed, dipoles=case.get("dipoles", 0),
                        integrator=case.get("integrator", 0), bounce=case.get("bounce"))
    sim = O.Sim(cfg)
    sim.init_particles(splitmix64_words(case["seed"]))
    return sim


def gpu_config(case, seed=0):
    a = case["args"]
    L = O.lib()
    return G.config(N=int(a["-N"]), r=int(a["-r"]), mpp=float(a["-m"]), dt=float(a["-d"]),
                    soft=float(a["-s"]), a=float(a["-a"]),
                    Te=L.or_strtold_to_double(a["-E"].encode()),
                    Ti=L.or_strtold_to_double(a["-I"].encode()),
                    x_drift=float(a.get("-x", "0")), settle=0.0, geometry=case["geometry"],
                    A=float(a.get("-A", "1")), theta=case["theta"], B=tuple(case["B"]), seed=seed,
                    dipoles=case.get("dipoles", 0), integrator=case.get("integrator", 0),
                    bounce=case.get("bounce"))


def gpu_ctx(case, parts, precision=G.POTG_FP64, seed=0, q=0, t=0, chi_b=0.0):
    ctx = G.Context([gpu_config(case, seed)], precision)
    ctx.upload(0, parts.view(G.PARTICLE_DT), q=q, t=t, chi_b=chi_b)
    return ctx


def domain_radius(case):
    return float(case["args"]["-r"]) * float(case["args"]["-m"])


def moved_mask(case, po, p0):
    """particles that moved more than 5% of the domain radius in the steps
    compared: the reinjected ones"""
    R = domain_radius(case)
    return np.sqrt((po["x"] - p0["x"]) ** 2 + (po["y"] - p0["y"]) ** 2 +
                   (po["z"] - p0["z"]) ** 2) > 0.05 * R


def assert_particles_match(case, po, pg, p0, x_tol=1e-12, v_tol=1e-10):
    """The float64 path against the oracle with the same semantics.  Bounds
    are relative to the domain radius (positions) and the largest velocity
    component (velocities).  Reinjected particles went through Newton
    inversions that stop at |p| < 1e-8 (prng.c:152, 211): libm and the
    device math library may stop one iteration apart, so those get a looser
    bound (1e-6 R, 1e-5 of the velocity scale)."""
    assert np.array_equal(pg["species"], po["species"])
    R = domain_radius(case)
    moved = moved_mask(case, po, p0)
    vscale = np.abs(po["vx"]).max()
    for f in ("x", "y", "z"):
        d = np.abs(pg[f] - po[f])
        assert d[~moved].max() <= x_tol * R, (f, d[~moved].max() / R)
        assert d[moved].max(initial=0) <= 1e-6 * R, f
    for f in ("vx", "vy", "vz"):
        d = np.abs(pg[f] - po[f])
        assert d[~moved].max() <= v_tol * vscale, (f, d[~moved].max() / vscale)
        assert d[moved].max(initial=0) <= 1e-5 * vscale, f


def assert_step_outputs_equal(oo, og):
    assert og["abs_q"] == oo.abs_q
    assert og["abs_e"] == oo.abs_e
    assert og["n_pushed"] == oo.n_pushed
    assert og["n_reinj_abs"] + og["n_reinj_out"] == oo.n_reinj_abs + oo.n_reinj_out
    assert og["n_wall_inward"] == oo.n_wall_inward


def step_both(case, steps=1, seed=77, precision=G.POTG_FP64, t0=0, chi_b=0.0, parts=None):
    """The oracle with the GPU's semantics (frozen leaf sources, the Philox
    reinjection stream) and a device context from the same state, the step
    counter set to t0 steps completed and the edge potential to chi_b;
    `steps` steps of each.  Returns (sim, ctx, oracle outputs, GPU outputs,
    start particles)."""
    sim = oracle_sim(case, inplace=0, rng_kind=O.RNG_PHILOX, philox_seed=seed)
    if parts is not None:
        sim.set_particles(parts)
    p0 = sim.particles()
    st = sim.state()
    sim.set_state(st["q"], st["m"], t0, chi_b, st["tot_e"])
    ctx = gpu_ctx(case, p0, precision=precision, seed=seed, q=st["q"], t=t0, chi_b=chi_b)
    oo, og = [], []
    for _ in range(steps):
        oo.append(sim.run_step())
        og.append(ctx.step()[0])
    return sim, ctx, oo, og, p0
