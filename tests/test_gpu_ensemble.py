"""The drop-in program on the CPU-runnable workload against the reference's
own runs of it, as statistics: 8192 electrons + 8192 ions around a grain of
0.1 lambda_D in a 10 lambda_D sphere, no B, no flow, theta 0.5, 2000 steps
(tests/golden/ensemble_16k: the reference, pot.c built by
oracle/build_ref.sh, run once per seed by make_golden.py --ensemble).

The GPU program runs the same seeds.  Its initial load is the reference's
bit for bit (test_host_init.py), but its reinjections draw from another
stream and its walk sums in another order, so runs of one seed part ways
within a few hundred steps; what must agree is the physics.  For each
quantity the mean over the seeds' runs is compared, program against
reference, in units of the combined standard error over seeds:

  - TE and TI averaged over steps 101-200, the span the reference records
    every step (pot.c:4364-4368); each run's own average also carries the
    lag-1 corrected standard error of pot.R:172-178, reported;
  - TE, TI, the grain charge DGQ and the electrons absorbed TOTE at step 2000.

The default program (float32 walk with compensated sums) is the one run.
A run in which chi_b bars a species from entering ends with the program's
reinjection error, where the reference's run never ends (pot.c:583-598);
such runs are left out of the program's set, as the reference's are of its
own."""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

from make_golden import splitmix64_words
from series_stats import across_runs, lag1_stats

from pot_amd import potg as G

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ENS = os.path.join(HERE, "golden", "ensemble_16k")
POT = os.path.join(os.path.dirname(G.LIB_PATH), "pot")
COLS = ("TIME", "DGQ", "TE", "TI", "TOTE")
Z_MAX = 3.5  # per comparison; with 6 comparisons a false alarm is ~0.3%


def _read_cols(path, cols=COLS):
    lines = open(path).read().split("\n")
    head = lines[0].split()
    idx = [head.index(c) for c in cols]
    rows = [[float(f.split()[k]) for k in idx] for f in lines[1:] if f.strip()]
    return np.array(rows)


def _summaries(tab):
    """per run: every-step span averages and the step-2000 record"""
    assert len(tab) == 201, len(tab)  # steps 1..200, then 2000
    span = tab[100:200]  # steps 101-200
    last = tab[200]
    return {"TE_101_200": span[:, 2], "TI_101_200": span[:, 3],
            "TE_2000": last[2], "TI_2000": last[3], "DGQ_2000": last[1], "TOTE_2000": last[4]}


@pytest.fixture(scope="module")
def runs(tmp_path_factory):
    meta = json.load(open(os.path.join(ENS, "ensemble.json")))
    ref, gpu = [], []
    for seed in meta["seeds"]:
        ref.append(_summaries(_read_cols(os.path.join(ENS, "seed_%d.dat" % seed))))
        d = tmp_path_factory.mktemp("ens%d" % seed)
        (d / "seed.bin").write_bytes(struct.pack("<16I", *splitmix64_words(seed)))
        argv = [POT]
        for k, v in meta["args"].items():
            argv += [k, v]
        argv += ["-i", str(meta["steps"]), "--theta", str(meta["theta"]),
                 "--seed-file", str(d / "seed.bin"), "--output-dir", str(d) + "/"]
        print("ensemble seed", seed, flush=True)
        r = subprocess.run(argv, capture_output=True, text=True, timeout=600)
        if r.returncode != 0 and "no admissible SCEPTIC reinjection" in r.stderr:
            # the case in which the reference's own run of a seed never ends
            # (ensemble.json, not_finished); left out here as those are there
            print("seed", seed, "left out:", r.stderr.strip().splitlines()[-1], flush=True)
            continue
        assert r.returncode == 0, r.stderr
        gpu.append(_summaries(_read_cols(str(d / "pot-macroscopic.dat"))))
    assert len(ref) >= 6 and len(gpu) >= 6, (len(ref), len(gpu))
    return ref, gpu


def _compare(ref, gpu, key, reduce=lambda v: v):
    a = across_runs([reduce(r[key]) for r in gpu])
    b = across_runs([reduce(r[key]) for r in ref])
    se = np.hypot(a["se"], b["se"])
    z = (a["mean"] - b["mean"]) / se if se > 0 else 0.0
    print("%-12s program %.5g +- %.3g, reference %.5g +- %.3g, z = %.2f" % (
        key, a["mean"], a["se"], b["mean"], b["se"], z))
    return z, a, b


@pytest.mark.parametrize("key", ["TE_101_200", "TI_101_200"])
def test_temperatures_over_the_recorded_span(runs, key):
    ref, gpu = runs
    z, a, b = _compare(ref, gpu, key, reduce=lambda v: float(np.mean(v)))
    assert abs(z) < Z_MAX, (key, a, b)
    # each run's own span average with the lag-1 correction (pot.R:172-178)
    for r in gpu + ref:
        s = lag1_stats(r[key])
        assert s["n"] == 100 and np.isfinite(s["se"]) and s["se"] > 0


@pytest.mark.parametrize("key", ["TE_2000", "TI_2000", "DGQ_2000", "TOTE_2000"])
def test_state_at_step_2000(runs, key):
    ref, gpu = runs
    z, a, b = _compare(ref, gpu, key)
    assert abs(z) < Z_MAX, (key, a, b)
