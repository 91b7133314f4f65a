"""The drop-in program `pot` end to end on the GPU (pot_amd/host/pot_main.c):
its output files against the oracle run with the GPU's semantics, and a
run resumed from its own state file against a run straight through.

Run the program with --seed-file: its initial load is the reference's
(bit-exact, test_host_init.py) and its reinjection stream is keyed by a mix
of the WELL512 record, restated below so the oracle draws the same numbers.
"""
import ctypes as C
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from gpu_common import big_case
from make_golden import splitmix64_words
from pot_amd import potg as G
from test_host_macro import MacroIn, _lib as macro_lib

pytestmark = pytest.mark.gpu

POT = os.path.join(os.path.dirname(G.LIB_PATH), "pot")
M64 = (1 << 64) - 1


def stream_key(prng_words):
    """the program's reinjection key: a splitmix64 chain over the 21 words of
    the WELL512 record (a, b, c, d, index, state[16])"""
    h = 0x9E3779B97F4A7C15
    for w in prng_words:
        z = h ^ w
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        h = z ^ (z >> 31)
    return h


def case_args(case):
    out = []
    for k, v in case["args"].items():
        out += [k, v]
    return out


def run_pot(tmp, args, name):
    d = tmp / name
    d.mkdir()
    r = subprocess.run([POT] + args + ["--output-dir", str(d) + "/"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return d


def oracle_macro_lines(case, seed, steps, integrator=0, bounce=None):
    sim = O.Sim(O.make_config(case["args"], geometry=0, theta=0.8, inplace=0))
    w = splitmix64_words(seed)
    sim.init_particles(w)
    pr = O.Prng()
    sim.L.or_get_prng(sim.h, C.byref(pr))
    key = stream_key([pr.a, pr.b, pr.c, pr.d, pr.index] + list(pr.state))
    sim = O.Sim(O.make_config(case["args"], geometry=0, theta=0.8, inplace=0,
                              rng_kind=O.RNG_PHILOX, philox_seed=key, integrator=integrator,
                              bounce=bounce))
    sim.init_particles(w)
    L = macro_lib()
    lines = []
    for _ in range(steps):
        sim.run_step()
        lines.append(format_oracle(L, sim, case))
    return lines


def format_oracle(L, sim, case):
    from test_host_macro import format_record
    return format_record(L, sim, dict(case, geometry=0))


def close_lines(a, b, rel=2e-5):
    fa, fb = a.split(), b.split()
    assert len(fa) == len(fb) == 35
    for k, (x, y) in enumerate(zip(fa, fb)):
        if k in (9, 34):  # grain charge and electrons absorbed: exact integers
            assert x == y, (k, x, y)
            continue
        vx, vy = float(x), float(y)
        assert abs(vx - vy) <= rel * max(abs(vx), abs(vy)) + 1e-300, (k, x, y)


def test_pot_program_matches_oracle(tmp_path):
    case = big_case(2001, seed=5)
    seed = 4242
    (tmp_path / "seed.bin").write_bytes(struct.pack("<16I", *splitmix64_words(seed)))
    steps = 12
    out = run_pot(tmp_path, case_args(case) + ["-i", str(steps), "--fp64", "--seed-file",
                                               str(tmp_path / "seed.bin"), "-f",
                                               str(tmp_path / "final.bin")], "run")
    got = open(out / "pot-macroscopic.dat").read().splitlines(True)
    assert got[0].startswith("TIME KE PE KI PI")
    assert len(got) == steps + 1
    want = oracle_macro_lines(case, seed, steps)
    for t in range(steps):
        close_lines(got[t + 1], want[t])
    info = open(out / "pot-run-init-info.txt").read().split("\t")
    assert info[1] == "2001" and info[-1].strip() == "0.8"
    assert "opening angle parameter: 0.8" in open(out / "pot-run-descrip.txt").read()
    hdr = open(tmp_path / "final.bin", "rb").readline().split()
    assert hdr[0] == b"C" and hdr[1] == b"1" and hdr[4] == str(steps).encode()


@pytest.mark.parametrize("opts,integrator,bounce,line", [
    (["--verlet"], 1, None, "numerical integrator: velocity Verlet"),
    (["--euler-richardson", "--bounce", "0.95"], 2, 0.95, "containment:          bouncing (c.o.e. 0.95)"),
])
def test_pot_program_other_integrators(tmp_path, opts, integrator, bounce, line):
    """The compile-time integrator and containment switches of the reference
    (pot.c:113-132) as options: macro records against the oracle, and the
    run description says which were used (pot.c:3751-3784)."""
    case = big_case(2001, seed=6)
    seed = 4343
    (tmp_path / "seed.bin").write_bytes(struct.pack("<16I", *splitmix64_words(seed)))
    steps = 8
    out = run_pot(tmp_path, case_args(case) + ["-i", str(steps), "--fp64", "--seed-file",
                                               str(tmp_path / "seed.bin")] + opts, "run")
    got = open(out / "pot-macroscopic.dat").read().splitlines(True)
    assert len(got) == steps + 1
    want = oracle_macro_lines(case, seed, steps, integrator=integrator, bounce=bounce)
    for t in range(steps):
        close_lines(got[t + 1], want[t])
    assert line in open(out / "pot-run-descrip.txt").read()


def test_resumed_run_continues_the_straight_run(tmp_path):
    """Resume after step 19: step 20 measures chi_b afresh (pot.c:2409-2418),
    so from there the particles follow the straight run exactly; only the
    grain mass, stored with 6 digits, differs (KG column)."""
    case = big_case(2001, seed=6)
    (tmp_path / "seed.bin").write_bytes(struct.pack("<16I", *splitmix64_words(99)))
    common = case_args(case) + ["--seed-file", str(tmp_path / "seed.bin")]
    a = run_pot(tmp_path, common + ["-i", "24"], "straight")
    run_pot(tmp_path, common + ["-i", "19", "-f", str(tmp_path / "s19.bin")], "first")
    b = run_pot(tmp_path, common + ["-i", "24", "-g", str(tmp_path / "s19.bin")], "resumed")
    la = open(a / "pot-macroscopic.dat").read().splitlines()
    lb = open(b / "pot-macroscopic.dat").read().splitlines()
    assert len(la) == 25 and len(lb) == 6  # header + steps 20..24
    for t in range(20, 25):
        ra, rb = la[t].split(), lb[t - 19].split()
        assert ra[:5] + ra[6:] == rb[:5] + rb[6:], t
