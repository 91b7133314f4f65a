"""The benchmark and parity workloads, expressed as reference command lines.

Each workload is a set of pot command-line values (strings, parsed the way
pot.c parses them) plus the switches the reference sets at compile time
(geometry, B, opening angle).  Plasma scales follow the reference's own
conventions:
  * lambda_D from n_e = (N - 1) / (2 V) (pot.c:3658-3666), so for N_e
    electrons in a sphere of R = k lambda_D:  R = 3 N_e e^2 / (4 pi eps0 kTe k^2);
  * dt = 0.01 / omega_pe with omega_pe = v_te / lambda_D, v_te = sqrt(kTe/me);
  * B from the electron gyroradius: B = me v_te / (e r_ge);
  * drift Mach number in units of sqrt(kTe/m_i);
  * softening 0.01 lambda_D; r = 1000 pixels and mpp = R / 1000.
"""
import math

import numpy as np

KB = 1.38065e-23
EPS0 = 8.85419e-12
QE = 1.6021766e-19
ME = 9.109383e-31
MIME = 1836.15


def _g(v):
    return "%.6g" % v


def plasma(n_e, r_over_ld, te=11600.0, te_over_ti=1.0, a_over_ld=0.1, rge_over_ld=None,
           b_dir=(0.0, 0.0, 1.0), mach=0.0, geometry=0, A=1.0, theta=0.5, semi_major=None):
    """Reference CLI values and compile-time switches for one grain run."""
    R = 3.0 * n_e * QE * QE / (4 * math.pi * EPS0 * KB * te * r_over_ld ** 2)
    ld = R / r_over_ld
    vte = math.sqrt(KB * te / ME)
    a = a_over_ld * ld
    if semi_major is not None:
        # the reference's a is the equal-volume radius: semi-major = a A^(2/3)
        a = semi_major * ld / A ** (2.0 / 3.0)
    args = {"-N": str(2 * n_e + 1), "-r": "1000", "-m": _g(R / 1000), "-a": _g(a),
            "-s": _g(0.01 * ld), "-d": _g(0.01 / (vte / ld)), "-E": _g(te),
            "-I": _g(te / te_over_ti), "-e": "0"}
    if mach:
        args["-x"] = _g(mach * math.sqrt(KB * te / (MIME * ME)))
    if geometry:
        args["-A"] = _g(A)
    B = (0.0, 0.0, 0.0)
    if rge_over_ld is not None:
        bmag = float(_g(ME * vte / (QE * rge_over_ld * ld)))
        n = np.array(b_dir, float)
        n /= np.linalg.norm(n)
        B = tuple(float(_g(bmag * c)) for c in n)
    return {"args": args, "geometry": geometry, "theta": theta, "B": B,
            "lambda_D": ld, "R": R}


def sphere_16k():
    """Plain CPU-runnable case: 8192 e + 8192 i, a = 0.1 lambda_D, R = 10 lambda_D,
    no B, no flow."""
    return plasma(8192, 10.0)


def sphere_2m_bz():
    """1M e + 1M i, R = 15 lambda_D, a = 0.1 lambda_D, B along z with
    r_ge = 0.5 lambda_D, Te/Ti = 1."""
    return plasma(1000000, 15.0, rge_over_ld=0.5)


def prolate_4m_tilted_b():
    """Prolate grain, aspect ratio 2, semi-major 0.2 lambda_D, 2M e + 2M i,
    R = 15 lambda_D, Te/Ti = 10, B tilted 30 degrees off the spheroid axis
    (x in the reference, SPH:1731-1737); r_ge = 0.5 lambda_D."""
    c30, s30 = math.cos(math.radians(30)), math.sin(math.radians(30))
    return plasma(2000000, 15.0, te_over_ti=10.0, rge_over_ld=0.5, b_dir=(c30, 0.0, s30),
                  geometry=1, A=2.0, semi_major=0.2)


def batched_sweep(n_inst=64, seed=2024):
    """The batched sweep: per-instance N log-uniform in [20k, 500k], a in
    [0.05, 0.5] lambda_D, r_ge/lambda_D in [0.2, 5], ion drift Mach in
    [0, 1.5] along B (x in the reference's drift convention), Te/Ti in {1, 10}."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_inst):
        n = int(round(math.exp(rng.uniform(math.log(2e4), math.log(5e5)))))
        n_e = n // 2
        out.append(plasma(n_e, 15.0, a_over_ld=float(rng.uniform(0.05, 0.5)),
                          rge_over_ld=float(rng.uniform(0.2, 5.0)), b_dir=(1.0, 0.0, 0.0),
                          mach=float(rng.uniform(0.0, 1.5)),
                          te_over_ti=float(rng.choice([1.0, 10.0]))))
    return out


def flowing_16m(theta=0.5):
    """8M e + 8M i, R = 20 lambda_D, Te/Ti = 10, ion drift Mach 1 along B,
    B along x (the reference's drift axis) with r_ge = lambda_D."""
    return plasma(8000000, 20.0, te_over_ti=10.0, rge_over_ld=1.0, b_dir=(1.0, 0.0, 0.0),
                  mach=1.0, theta=theta)


def potg_config(w, seed=1):
    """potg_config for a workload (pot_amd.potg.Config)."""
    from pot_amd import potg as G
    a = w["args"]
    return G.config(N=int(a["-N"]), r=int(a["-r"]), mpp=float(a["-m"]), dt=float(a["-d"]),
                    soft=float(a["-s"]), a=float(a["-a"]), Te=float(a["-E"]),
                    Ti=float(a["-I"]), x_drift=float(a.get("-x", "0")),
                    settle=float(a.get("-e", "0")), geometry=w["geometry"],
                    A=float(a.get("-A", "1")), theta=w["theta"], B=w["B"], seed=seed)
