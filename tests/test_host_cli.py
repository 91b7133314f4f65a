"""The drop-in programs' command line (pot_amd/host/pot_main.c) on the CPU:
option parsing, settings listing, argument errors and the state-file
read-out (pot.c:4555-4961), none of which touches the GPU."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from gpu_common import CASES, GOLD
from pot_amd import potg as G

BUILD = os.path.dirname(G.LIB_PATH)
POT = os.path.join(BUILD, "pot")
POT_SPH = os.path.join(BUILD, "pot-spheroid")


def run(prog, *args, cwd=None):
    return subprocess.run([prog] + [str(a) for a in args], capture_output=True, text=True,
                          cwd=cwd, timeout=60)


def test_settings_listing_shows_defaults_and_options():
    r = run(POT, "-l")
    assert r.returncode == 0
    assert r.stdout == ("Settings: %s -N 150001 -r 1000 -d 1e-12 -m 4e-07 -s 2e-08 -E 220 -I 220 "
                        "-x 0 -a 2.5e-06 -e 5e-07\n" % POT)
    r = run(POT, "-N", 2001, "-i", 7, "-d", "2.5e-16", "-f", "out.bin", "-w", 3, "-l")
    assert r.returncode == 0
    assert " -N 2001 " in r.stdout and " -d 2.5e-16 " in r.stdout
    assert r.stdout.rstrip().endswith("-i 7 -f out.bin -w 3")
    r = run(POT_SPH, "-l")
    assert r.returncode == 0
    assert (" -r 654 -A 5 -d 1e-12 -m 4e-07 -s 4e-10 -E 11600 -I 11600 -x 0 -a 2.5e-06 -e 0\n"
            in r.stdout)


@pytest.mark.parametrize("args,msg", [
    (["-N", "0"], "Number of particles must be a valid nonzero integer."),
    (["-d", "-1"], "Time step must be positive."),
    (["-a", "0"], "Grain radius must be positive."),
    (["-s", "-1e-9"], "Softening parameter must be non-negative."),
    (["-i", "x"], "Number of time steps to run for must be a valid nonzero integer."),
    (["-r", "0"], "Simulation radius in pixels must be a valid nonzero integer."),
])
def test_bad_values_are_refused(args, msg):
    r = run(POT, *args)
    assert r.returncode == 1
    assert msg in r.stderr


def test_unknown_option_and_missing_value_show_usage():
    for args in (["-q"], ["-N"], ["stray"], ["--nonsense", "1"], ["-A", "2"]):
        r = run(POT, *args)
        assert r.returncode == 1
        assert r.stderr.startswith("Usage: ")


def test_spheroid_aspect_ratio_checked_against_shape():
    assert "Prolate aspect ratio must be greater than 1." in run(POT_SPH, "-A", "0.5").stderr
    assert "Oblate aspect ratio must be less than 1." in run(POT_SPH, "--oblate", "-A", "2").stderr
    assert run(POT_SPH, "-A", "0.5", "--oblate", "-l").returncode == 0


def test_grain_must_fit_in_domain():
    r = run(POT, "-r", 10, "-m", "1e-7", "-a", "2e-6")
    assert r.returncode == 1
    assert "Grain radius (2e-06 m) must be less than simulation region radius (1e-06 m)." in r.stderr


@pytest.mark.parametrize("case", CASES[:2], ids=[c["name"] for c in CASES[:2]])
def test_read_out_of_a_reference_state_file(case):
    N = int(case["args"]["-N"])
    path = os.path.join(GOLD, case["name"], "state_1.bin")
    r = run(POT, "-N", N, "-R", path)
    assert r.returncode == 0
    lines = r.stdout.splitlines()
    assert lines[0] == "NO S X Y Z VX VY VZ"
    assert len(lines) == N + 1
    _, _, parts = O.read_state_file(path, N)
    for k in (0, 1, N // 2, N - 1):
        p = parts[k]
        want = '%u "%s" %s' % (k, "die"[p["species"]], " ".join(
            "%g" % v for v in (p["x"], p["y"], p["z"], p["vx"], p["vy"], p["vz"])))
        assert lines[k + 1] == want


def test_state_file_with_other_worker_count_is_refused(tmp_path):
    case = CASES[0]
    N = int(case["args"]["-N"])
    raw = open(os.path.join(GOLD, case["name"], "state_1.bin"), "rb").read()
    bad = tmp_path / "s.bin"
    bad.write_bytes(b"C 4 " + raw[4:])
    r = run(POT, "-N", N, "-R", str(bad))
    assert r.returncode == 1
    assert "conflicting worker process counts (4 vs. 1)" in r.stderr
