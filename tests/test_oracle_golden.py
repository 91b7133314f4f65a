"""The oracle against runs of the reference itself.

tests/golden/*/state_{1,K}.bin were written by the reference binary built
from its own sources (oracle/build_ref.sh, tests/golden/make_golden.py).
The oracle, in reference mode (WELL512 stream, in-place half drift, long
double arithmetic), must reproduce those states bit for bit: this pins the
restated WELL512 and samplers, init_particles, build_tree and moments, the
field walk, the Boris push, grain crossing and tally, SCEPTIC reinjection
and chi_b (pot.c:496-2934, SPH:1689-2540, prng.c, coord.c).
"""
import json
import os
import tempfile

import numpy as np
import pytest

import oracle_lib as O
from make_golden import splitmix64_words

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))


def run_oracle(case, steps):
    cfg = O.make_config(case["args"], geometry=case["geometry"], theta=case["theta"],
                        B=tuple(case["B"]), dipoles=case.get("dipoles", 0),
                        integrator=case.get("integrator", 0), bounce=case.get("bounce"))
    sim = O.Sim(cfg)
    sim.init_particles(splitmix64_words(case["seed"]))
    for _ in range(steps):
        sim.run_step()
    return sim


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
@pytest.mark.parametrize("which", ["state_1.bin", "state_K.bin"])
def test_oracle_reproduces_reference_state(case, which):
    N = int(case["args"]["-N"])
    steps = 1 if which == "state_1.bin" else case["steps"]
    hdr_ref, prng_ref, parts_ref = O.read_state_file(os.path.join(GOLD, case["name"], which), N)
    sim = run_oracle(case, steps)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "s.bin")
        sim.write_state(p)
        hdr, prng, parts = O.read_state_file(p, N)
    assert hdr == hdr_ref
    assert np.array_equal(prng, prng_ref)
    for f in ("x", "y", "z", "vx", "vy", "vz", "species"):
        a, b = parts[f], parts_ref[f]
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, "%s differs at %s (%r vs %r)" % (f, bad[:5], a[bad[:5]], b[bad[:5]])
