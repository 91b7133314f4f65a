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
