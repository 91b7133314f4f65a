#!/usr/bin/env python3
"""Write the reference-run fixtures under tests/golden/.

Runs the reference treecode, built from its own sources by
oracle/build_ref.sh into oracle/_ref/, on small seeded cases and keeps what
the reference itself writes: the format-C state file after 1 step and
after K steps (pot.c:3816-3901) and pot-macroscopic.dat (pot.c:3247-3510).
The PRNG seed is the 64-byte file ./seed.bin (RANDOMNESS_SOURCE, pot.c:192)
expanded from an integer by splitmix64, so a test can rebuild it.

Only this container has the reference; the fixtures travel, the reference
does not.  Usage: python tests/golden/make_golden.py [--ensemble | --paper]
(--ensemble: only the 2000-step runs of the CPU-runnable workload; --paper:
the grain-charge series of the reference's published run, below)
"""
import json
import math
import os
import shutil
import struct
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(REPO, "oracle", "_ref")

KB = 1.38065e-23
EPS0 = 8.85419e-12
QE = 1.6021766e-19
ME = 9.109383e-31


def splitmix64_words(seed, n=16):
    """n little-endian u32 words from a splitmix64 stream (Steele et al. 2014)."""
    out = []
    x = seed & 0xFFFFFFFFFFFFFFFF
    while len(out) < n:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        z ^= z >> 31
        out.append(z & 0xFFFFFFFF)
        out.append(z >> 32)
    return out[:n]


def seed_bytes(seed):
    return struct.pack("<16I", *splitmix64_words(seed))


def plasma_args(N, r_over_ld, te, ti, a_over_ld, soft_over_ld, dt_frac, r_px=1000):
    """CLI values for a domain of R/lambda_D with N-1 plasma particles.

    lambda_D follows the reference's convention n_e = (N-1)/(2V)
    (pot.c:3658-3666); dt is dt_frac / omega_pe."""
    n_e_count = (N - 1) // 2
    R = 3.0 * n_e_count * QE * QE / (4 * math.pi * EPS0 * KB * te * r_over_ld ** 2)
    ld = R / r_over_ld
    wpe = math.sqrt(KB * te / ME) / ld
    fmt = lambda v: "%.6g" % v
    return {
        "-N": str(N), "-r": str(r_px), "-m": fmt(R / r_px), "-a": fmt(a_over_ld * ld),
        "-s": fmt(soft_over_ld * ld), "-d": fmt(dt_frac / wpe), "-E": fmt(te),
        "-I": fmt(ti), "-e": "0",
    }


def cases():
    base = plasma_args(601, 5.0, 11600.0, 11600.0, 0.5, 0.01, 0.01)
    cs = []
    cs.append({"name": "sphere", "bin": "pot_sphere", "geometry": 0, "theta": 0.5,
               "B": [0.0, 0.0, 0.0], "args": dict(base), "steps": 100, "seed": 1})
    cs.append({"name": "sphere_bz", "bin": "pot_sphere_bz", "geometry": 0, "theta": 0.5,
               "B": [0.0, 0.0, 50.0], "args": dict(base), "steps": 100, "seed": 2})
    drift = dict(base)
    drift["-x"] = "%.6g" % math.sqrt(KB * 11600.0 / (1836.15 * ME))
    cs.append({"name": "sphere_bx_drift", "bin": "pot_sphere_bx", "geometry": 0,
               "theta": 0.5, "B": [30.0, 0.0, 0.0], "args": drift, "steps": 100, "seed": 3})
    cs.append({"name": "sphere_t08", "bin": "pot_sphere_t08", "geometry": 0, "theta": 0.8,
               "B": [0.0, 0.0, 0.0], "args": dict(base), "steps": 50, "seed": 4})
    pro = plasma_args(601, 5.0, 11600.0, 1160.0, 0.8, 0.01, 0.01)
    pro["-A"] = "2"
    cs.append({"name": "prolate", "bin": "pot_prolate", "geometry": 1, "theta": 0.5,
               "B": [43.30127018922193, 0.0, 25.0], "args": pro, "steps": 150, "seed": 5})
    obl = plasma_args(601, 5.0, 11600.0, 11600.0, 0.5, 0.01, 0.01)
    obl["-A"] = "0.5"
    cs.append({"name": "oblate", "bin": "pot_oblate", "geometry": 2, "theta": 0.5,
               "B": [0.0, 0.0, 0.0], "args": obl, "steps": 60, "seed": 6})
    cs.append({"name": "sphere_dipoles", "bin": "pot_sphere_dip", "geometry": 0, "theta": 0.5,
               "B": [0.0, 0.0, 20.0], "args": dict(base), "steps": 60, "seed": 7,
               "dipoles": 1})
    # the other integrators (1 velocity Verlet, 2 Euler-Richardson) and
    # bouncing containment (pot.c:113-132)
    cs.append({"name": "sphere_verlet_bz", "bin": "pot_sphere_verlet_bz", "geometry": 0,
               "theta": 0.5, "B": [0.0, 0.0, 20.0], "args": dict(base), "steps": 60, "seed": 8,
               "integrator": 1})
    cs.append({"name": "sphere_er_bounce_bx", "bin": "pot_sphere_er_bounce_bx", "geometry": 0,
               "theta": 0.5, "B": [30.0, 0.0, 0.0], "args": dict(base), "steps": 60, "seed": 9,
               "integrator": 2, "bounce": 0.95})
    cs.append({"name": "sphere_bounce", "bin": "pot_sphere_bounce", "geometry": 0, "theta": 0.5,
               "B": [0.0, 0.0, 0.0], "args": dict(base), "steps": 60, "seed": 10,
               "bounce": 0.95})
    cs.append({"name": "prolate_verlet", "bin": "pot_prolate_verlet", "geometry": 1,
               "theta": 0.5, "B": [0.0, 0.0, 0.0], "args": dict(pro), "steps": 60, "seed": 11,
               "integrator": 1})
    return cs


def run_case(c, steps, workdir, timeout=600):
    os.makedirs(os.path.join(workdir, "output"), exist_ok=True)
    with open(os.path.join(workdir, "seed.bin"), "wb") as f:
        f.write(seed_bytes(c["seed"]))
    argv = [os.path.join(REF_BIN, c["bin"])]
    for k, v in c["args"].items():
        argv += [k, v]
    argv += ["-i", str(steps), "-f", "state.bin"]
    subprocess.run(argv, cwd=workdir, check=True, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL, timeout=timeout)
    return workdir


# Ensemble of the CPU-runnable workload (8192 e + 8192 i, a = 0.1 lambda_D,
# R = 10 lambda_D, no B, no flow, theta 0.5, dt = 0.01/omega_pe; the
# sphere_16k workload of pot_amd/workloads.py), 2000 steps per seed.  The
# reference records every step up to t = 200 and then every 2000th
# (pot.c:4364-4368 with settle 0); the columns kept are TIME, DGQ, TE, TI
# and TOTE of its pot-macroscopic.dat (pot.c:3247-3510).
ENSEMBLE_SEEDS = tuple(range(1001, 1011))
ENSEMBLE_STEPS = 2000
ENSEMBLE_TIMEOUT = 900  # s; the finished runs take about 6 minutes
ENSEMBLE_COLS = ("TIME", "DGQ", "TE", "TI", "TOTE")


def ensemble_args():
    sys.path.insert(0, REPO)
    from pot_amd import workloads as W
    return dict(W.sphere_16k()["args"])


def run_ensemble_seed(seed, dest):
    """True when the reference finished the run.  A run can fail to finish:
    when chi_b bars electrons from entering, sceptically_reinject recurses
    without end (pot.c:583-598)."""
    out = os.path.join(dest, "seed_%d.dat" % seed)
    if os.path.exists(out):
        return True
    c = {"bin": "pot_sphere", "args": ensemble_args(), "seed": seed}
    wd = tempfile.mkdtemp(prefix="potens_")
    try:
        try:
            run_case(c, ENSEMBLE_STEPS, wd, timeout=ENSEMBLE_TIMEOUT)
        except subprocess.TimeoutExpired:
            return False
        lines = open(os.path.join(wd, "output", "pot-macroscopic.dat")).read().split("\n")
        head = lines[0].split()
        cols = [head.index(k) for k in ENSEMBLE_COLS]
        rows = [" ".join(ENSEMBLE_COLS)]
        for ln in lines[1:]:
            f = ln.split()
            if f:
                rows.append(" ".join(f[k] for k in cols))
        with open(out, "w") as fh:
            fh.write("\n".join(rows) + "\n")
        return True
    finally:
        shutil.rmtree(wd)


def ensemble():
    """The reference's runs of the CPU-runnable workload, one per seed, in
    parallel (one process each); runs already written are kept."""
    from concurrent.futures import ThreadPoolExecutor
    dest = os.path.join(HERE, "ensemble_16k")
    os.makedirs(dest, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(ENSEMBLE_SEEDS))) as ex:
        done = list(ex.map(lambda s: run_ensemble_seed(s, dest), ENSEMBLE_SEEDS))
    with open(os.path.join(dest, "ensemble.json"), "w") as fh:
        json.dump({"args": ensemble_args(), "bin": "pot_sphere", "theta": 0.5,
                   "steps": ENSEMBLE_STEPS, "columns": list(ENSEMBLE_COLS),
                   "seeds": [s for s, d in zip(ENSEMBLE_SEEDS, done) if d],
                   "not_finished": {str(s): "no end within %d s: sceptically_reinject kept "
                                            "retrying (pot.c:583-598)" % ENSEMBLE_TIMEOUT
                                    for s, d in zip(ENSEMBLE_SEEDS, done) if not d}},
                  fh, indent=1)
    print("wrote", dest)


# The grain-charge series of the reference's published run with
# particle-particle forces in a spherical domain (output/2016 TH PPCF
# paper/np-with-sph-macro.dat; gasma.R:119-127 plots it): TIME and DGQ only.
PAPER_SRC = os.path.join("/root/reference", "output", "2016 TH PPCF paper",
                         "np-with-sph-macro.dat")


def paper():
    dest = os.path.join(HERE, "paper_2016")
    os.makedirs(dest, exist_ok=True)
    lines = open(PAPER_SRC).read().split("\n")
    head = lines[0].split()
    it, iq = head.index("TIME"), head.index("DGQ")
    out = ["TIME DGQ"]
    for ln in lines[1:]:
        f = ln.split()
        if f:
            out.append("%s %s" % (f[it], f[iq]))
    with open(os.path.join(dest, "np_with_sph_dgq.dat"), "w") as fh:
        fh.write("\n".join(out) + "\n")
    print("wrote", dest, len(out) - 1, "records")


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--paper":
        paper()
        return
    if not os.path.isdir(REF_BIN):
        sys.exit("build the reference variants first: oracle/build_ref.sh")
    if len(sys.argv) > 1 and sys.argv[1] == "--ensemble":
        ensemble()
        return
    all_cases = cases()
    for c in all_cases:
        dest = os.path.join(HERE, c["name"])
        os.makedirs(dest, exist_ok=True)
        for steps, tag in ((1, "state_1.bin"), (c["steps"], "state_K.bin")):
            wd = tempfile.mkdtemp(prefix="potgold_")
            try:
                run_case(c, steps, wd)
                shutil.copy(os.path.join(wd, "state.bin"), os.path.join(dest, tag))
                if tag == "state_K.bin":
                    shutil.copy(os.path.join(wd, "output", "pot-macroscopic.dat"),
                                os.path.join(dest, "macro.dat"))
            finally:
                shutil.rmtree(wd)
        print("wrote", dest)
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump(all_cases, f, indent=1)


if __name__ == "__main__":
    main()
