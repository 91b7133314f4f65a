"""bench.py's command line on the CPU: it measures one GPU only and says so
for more, before it touches a device."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_paper_stats_of_a_saved_series(tmp_path):
    """The paper runs' statistics from a saved series, on the CPU: runs whose
    charge is the reference's published series shifted by +1, -1, +2 and -2
    elementary charges have the reference's mean over the same window, a
    standard error across runs of sqrt(10/3)/2, and a difference of zero."""
    import json
    import numpy as np
    ref = np.loadtxt(os.path.join(REPO, "tests", "golden", "paper_2016", "np_with_sph_dgq.dat"),
                     skiprows=1)
    # one record every 2 ns from 2 to 2.5 us, as the runs record
    t = np.round(np.arange(2000, 2502, 2) * 1e-9, 12)
    idx = [int(np.argmin(np.abs(ref[:, 0] - x))) for x in t]
    assert np.allclose(ref[idx, 0], t, rtol=0, atol=1e-15)
    q = ref[idx, 1]
    with open(tmp_path / "series.csv", "w") as f:
        f.write("step,time,q0,q1,q2,q3\n")
        for k, (x, v) in enumerate(zip(t, q)):
            f.write("%d,%.6e,%d,%d,%d,%d\n" % (int(round(x / 1e-12)), x, v + 1, v - 1, v + 2, v - 2))
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--mode", "paper-stats",
                        "--state-dir", str(tmp_path), "--t-eq", "2e-6"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    same = d["reference_same_window"]
    assert same["n"] == len(t)
    assert abs(d["mean"] - q.mean()) < 1e-9 and abs(same["mean"] - q.mean()) < 1e-9
    assert abs(same["difference"]) < 1e-9
    assert abs(d["se_across_runs"] - np.sqrt(10.0 / 3.0) / 2.0) < 1e-12
    assert d["reference_all_after_t_eq"]["n"] > same["n"]
