"""Parity of the HIP path with the oracle on the same seeded snapshots.

Tree: bit-exact (node set, particle per leaf, q, aq, centre of charge,
box size) against the oracle's build_tree (pot.c:1457-1643).
FP64 walk: the same interaction set (exact counts) and the same float64
operations as the oracle's float64-accumulating walk; the results must
agree to the last bit or, at worst, to 1e-13 relative.
Mixed walk: float32 with compensated sums against the float64 oracle
walk; tolerance stated per test.
"""
import numpy as np
import pytest

import oracle_lib as O
from gpu_common import CASES, big_case, gpu_ctx, oracle_sim

pytestmark = pytest.mark.gpu

from pot_amd import potg as G

SMALL = [c for c in CASES]
BIG = [big_case(20001), big_case(20001, geometry=1, A=2.0, seed=12),
       big_case(20001, geometry=2, A=0.5, seed=13),
       dict(big_case(20001, seed=14), name="big20001_dipoles", dipoles=1)]
ALL = SMALL + BIG
IDS = [c["name"] for c in ALL]


def canon_equal(a, b):
    for f in ("depth", "is_leaf", "q", "aq", "nchild"):
        assert np.array_equal(a[f], b[f]), f
    leaf = a["is_leaf"] == 1
    assert np.array_equal(a["part"][leaf], b["part"][leaf])
    assert np.array_equal(a["c"], b["c"]), "centre of charge"


@pytest.mark.parametrize("case", ALL, ids=IDS)
def test_tree_bit_exact(case):
    sim = oracle_sim(case)
    parts = sim.particles()
    sim.build_tree()
    ref = sim.tree_canon()
    ctx = gpu_ctx(case, parts)
    ctx.build_tree()
    got = ctx.tree_canon(0)
    assert len(got["depth"]) == len(ref["depth"])
    canon_equal(got, ref)
    internal = ref["is_leaf"] == 0
    size2 = (ref["x1x2"][:, 1] - ref["x1x2"][:, 0]) ** 2
    assert np.array_equal(got["size2"][internal], size2[internal])
    if case.get("dipoles"):
        # the dipole fold of update_node_moments (pot.c:1300-1307), bit for bit
        assert np.array_equal(ctx.tree_dipoles(0), sim.tree_dipoles())


@pytest.mark.parametrize("N,geometry", [(4096, 0), (4097, 0), (12289, 0), (4097, 1), (8194, 2)])
def test_tree_bit_exact_at_sort_tile_edges(N, geometry):
    """The radix sort works on tiles of 4096 keys: trees of exactly one tile,
    one key past it and three tiles and one key, bit for bit; with a
    spheroid the grain's key, which sorts last, is alone in the last tile
    (4097) or shares it with one particle (8194)."""
    case = big_case(N, geometry=geometry, A=2.0 if geometry == 1 else 0.5, seed=N + geometry)
    sim = oracle_sim(case)
    parts = sim.particles()
    sim.build_tree()
    ctx = gpu_ctx(case, parts, precision=G.POTG_MIXED)
    ctx.build_tree()
    got, ref = ctx.tree_canon(0), sim.tree_canon()
    assert len(got["depth"]) == len(ref["depth"])
    canon_equal(got, ref)


@pytest.mark.parametrize("case", ALL, ids=IDS)
def test_walk_fp64_matches_oracle(case):
    sim = oracle_sim(case)
    parts = sim.particles()
    sim.build_tree()
    ctx = gpu_ctx(case, parts)
    ctx.build_tree()
    idx = np.arange(1, len(parts), dtype=np.uint32)
    xyz = np.stack([parts["x"][idx], parts["y"][idx], parts["z"][idx]], 1)
    E, cnt = ctx.field_at(0, xyz, exclude=idx, counts=True)
    Eo = np.zeros_like(E)
    cnto = np.zeros_like(cnt, dtype=np.int64)
    for k, i in enumerate(idx):
        st = O.WalkStats()
        Eo[k] = sim.field_at(int(i), xyz[k, 0], xyz[k, 1], xyz[k, 2], 1, st)
        cnto[k] = (st.n_cell, st.n_leaf, st.n_open)
    assert np.array_equal(cnt.astype(np.int64), cnto), "interaction counts differ"
    mag = np.linalg.norm(Eo, axis=1)
    rel = np.linalg.norm(E - Eo, axis=1) / mag
    assert rel.max() <= 1e-13, rel.max()


@pytest.mark.parametrize("case", ALL, ids=IDS)
def test_walk_mixed_close_to_fp64(case):
    sim = oracle_sim(case)
    parts = sim.particles()
    sim.build_tree()
    ctx = gpu_ctx(case, parts, precision=G.POTG_MIXED)
    ctx.build_tree()
    idx = np.arange(1, len(parts), dtype=np.uint32)
    xyz = np.stack([parts["x"][idx], parts["y"][idx], parts["z"][idx]], 1)
    E = ctx.field_at(0, xyz, exclude=idx)
    Eo = np.array([sim.field_at(int(i), *xyz[k], 1) for k, i in enumerate(idx)])
    rel = np.linalg.norm(E - Eo, axis=1) / np.linalg.norm(Eo, axis=1)
    # float32 walk: positions carry ~6e-8 R; MAC decisions may flip at the
    # threshold, which moves a few particles by the truncation error
    assert np.median(rel) < 1e-5, np.median(rel)
    assert np.percentile(rel, 99) < 2e-3, np.percentile(rel, 99)


@pytest.mark.parametrize("case", ALL, ids=IDS)
def test_direct_sum_matches_oracle(case):
    sim = oracle_sim(case)
    parts = sim.particles()
    sim.build_tree()
    ctx = gpu_ctx(case, parts)
    idx = np.arange(1, len(parts), 7, dtype=np.uint32)
    E = ctx.direct_sum(0, idx)
    Eo = np.array([sim.direct_sum(int(i), parts["x"][i], parts["y"][i], parts["z"][i]) for i in idx])
    rel = np.linalg.norm(E - Eo, axis=1) / np.linalg.norm(Eo, axis=1)
    assert rel.max() < 1e-12, rel.max()


def _step_both(case, steps=1, seed=77):
    sim = oracle_sim(case, inplace=0, rng_kind=O.RNG_PHILOX, philox_seed=seed)
    parts = sim.particles()
    ctx = gpu_ctx(case, parts, seed=seed)
    outs_o, outs_g = [], []
    for _ in range(steps):
        outs_o.append(sim.run_step())
        outs_g.append(ctx.step()[0])
    return sim, ctx, outs_o, outs_g


@pytest.mark.parametrize("case", SMALL, ids=[c["name"] for c in SMALL])
def test_step_matches_oracle_frozen_philox(case):
    """One full step (tree, walk, Boris, crossing, tally, reinjection) of the
    GPU FP64 path against the oracle run with the GPU's semantics: frozen
    leaf sources and the same Philox reinjection stream."""
    sim, ctx, oo, og = _step_both(case, steps=1)
    po = sim.particles()
    pg, gr = ctx.download(0)
    assert og[0]["abs_q"] == oo[0].abs_q
    assert og[0]["abs_e"] == oo[0].abs_e
    assert og[0]["n_pushed"] == oo[0].n_pushed
    assert og[0]["n_reinj_abs"] + og[0]["n_reinj_out"] == oo[0].n_reinj_abs + oo[0].n_reinj_out
    assert np.array_equal(pg["species"], po["species"])
    R = float(case["args"]["-r"]) * float(case["args"]["-m"])
    # reinjected particles went through Newton inversions that stop at
    # |p| < 1e-8 (prng.c:152, 211): libm and the device math library may
    # stop one iteration apart, so those particles get a looser bound
    p0 = oracle_sim(case).particles()
    moved = np.sqrt((po["x"] - p0["x"]) ** 2 + (po["y"] - p0["y"]) ** 2 +
                    (po["z"] - p0["z"]) ** 2) > 0.05 * R
    vscale = np.abs(po["vx"]).max()
    for f in ("x", "y", "z"):
        d = np.abs(pg[f] - po[f])
        assert d[~moved].max() <= 1e-12 * R, f
        assert d[moved].max(initial=0) <= 1e-6 * R, f
    for f in ("vx", "vy", "vz"):
        d = np.abs(pg[f] - po[f])
        assert d[~moved].max() <= 1e-10 * vscale, f
        assert d[moved].max(initial=0) <= 1e-5 * vscale, f
    assert gr["q"] == sim.state()["q"]


@pytest.mark.parametrize("case", SMALL, ids=[c["name"] for c in SMALL])
def test_step_potential_energies_match_oracle(case):
    """compute_epe steps: the potential each particle feels at its
    half-drifted position (phi_particle_feels, pot.c:1934-2018) summed per
    species, and the grain's with its updated charge (pot.c:2852-2863),
    against the oracle (whose sums the reference's macro files pin).  The
    oracle adds in long double in index order, the GPU in float64 in a
    fixed tree order: 1e-11 relative."""
    import ctypes as C
    sim = oracle_sim(case, inplace=0, rng_kind=O.RNG_PHILOX, philox_seed=77)
    ctx = gpu_ctx(case, sim.particles(), seed=77)
    for step in range(3):
        sim.run_step()  # every step up to 200 computes the energies
        ctx.step(compute_epe=1)
        epe = (C.c_longdouble * 3)()
        sim.L.or_get_state(sim.h, None, None, None, None, None, epe)
        _, gr = ctx.download(0)
        for s in range(3):
            want = float(epe[s])
            if want == 0.0:
                assert gr["epe"][s] == 0.0, (step, s)
            else:
                assert abs(gr["epe"][s] - want) <= 1e-11 * abs(want), (step, s, gr["epe"][s], want)


def test_tuning_takes_no_step_and_is_reproducible():
    """potg_tune times the coming step's walk without taking the step: a
    context that tunes and then steps matches one that was given the same
    launch parameters directly, bit for bit, and the step counter and work
    counters are untouched by the tuning."""
    case = big_case(20001, seed=23)
    parts = oracle_sim(case).particles()
    for prec in (G.POTG_MIXED, G.POTG_FP64):
        a = gpu_ctx(case, parts, precision=prec, seed=3)
        b = gpu_ctx(case, parts, precision=prec, seed=3)
        a.step()
        b.step()
        before = a.stats()
        t = a.tune()
        assert t["walk_blocks"] > 0 and t["walk_ms"] > 0
        assert t["group"] in (8, 16, 64)
        after = a.stats()
        for k in ("steps", "n_cell", "n_leaf", "n_open", "n_range_loads"):
            assert after[k] == before[k], k
        b.set_tuning(t["walk_blocks"], t["group"], t["variant"], t["sort_tile"])
        for _ in range(3):
            a.step()
            b.step()
        pa, _ = a.download(0)
        pb, _ = b.download(0)
        for f in ("x", "y", "z", "vx", "vy", "vz", "species"):
            assert np.array_equal(pa[f], pb[f]), (prec, f)
        assert a.stats()["steps"] == 4


@pytest.mark.parametrize("precision,group", [(G.POTG_FP64, 64), (G.POTG_MIXED, 64),
                                             (G.POTG_MIXED, 16), (G.POTG_MIXED_GROUP, 64)],
                         ids=["fp64", "mixed64", "mixed16", "mixed_group"])
@pytest.mark.parametrize("dipoles", [0, 1], ids=["monopole", "dipoles"])
def test_counting_off_changes_only_the_counters(precision, group, dipoles):
    """potg_set_counting(0) drops the interaction counters of the step walks
    and nothing else: the same steps end in the same state bit for bit, the
    step and target counts are kept, and point queries still count."""
    case = dict(big_case(20001, seed=31), dipoles=dipoles)
    parts = oracle_sim(case).particles()
    a = gpu_ctx(case, parts, precision=precision, seed=5)
    b = gpu_ctx(case, parts, precision=precision, seed=5)
    for c in (a, b):
        c.set_tuning(2048, group)
    b.set_counting(False)
    for _ in range(3):
        a.step()
        b.step()
    pa, ga = a.download(0)
    pb, gb = b.download(0)
    assert np.array_equal(pa.view(np.uint8), pb.view(np.uint8))
    assert ga == gb
    sa, sb = a.stats(), b.stats()
    assert sa["n_cell"] > 0 and sa["n_leaf"] > 0 and sa["n_open"] > 0
    for k in ("n_cell", "n_leaf", "n_open", "n_range_loads"):
        assert sb[k] == 0, k
    for k in ("steps", "n_targets"):
        assert sa[k] == sb[k], k
    # a graph replay after switching counting back on counts again
    b.set_counting(True)
    b.run(2)
    assert b.stats()["n_cell"] > 0
    idx = np.arange(1, 2001, 7, dtype=np.uint32)
    xyz = np.stack([pb["x"][idx], pb["y"][idx], pb["z"][idx]], 1)
    b.set_counting(False)
    _, cnt = b.field_at(0, xyz, exclude=idx, counts=True)
    assert cnt[:, 1].min() > 0


@pytest.mark.parametrize("group,blocks", [(64, 2048), (64, 300), (8, 2048), (16, 2048)])
@pytest.mark.parametrize("case", [BIG[0], BIG[3]], ids=["monopole", "dipoles"])
def test_mixed_walk_shapes_agree(case, group, blocks):
    """Every walk shape (group width, grid size) gives the same
    interaction sets as the float64 walk (counts equal but for MAC decisions
    float32 flips at the threshold) and the same field to float32 accuracy."""
    sim = oracle_sim(case)
    parts = sim.particles()
    ctx64 = gpu_ctx(case, parts)
    ctx64.build_tree()
    ctx = gpu_ctx(case, parts, precision=G.POTG_MIXED)
    ctx.set_tuning(blocks, group)
    ctx.build_tree()
    idx = np.arange(1, len(parts), 3, dtype=np.uint32)
    xyz = np.stack([parts["x"][idx], parts["y"][idx], parts["z"][idx]], 1)
    E64, c64 = ctx64.field_at(0, xyz, exclude=idx, counts=True)
    E32, c32 = ctx.field_at(0, xyz, exclude=idx, counts=True)
    same = np.all(c32 == c64, axis=1)
    assert same.mean() > 0.99  # float32 MAC decisions may flip at the threshold
    rel = np.linalg.norm(E32 - E64, axis=1) / np.linalg.norm(E64, axis=1)
    assert np.median(rel) < 1e-5


def _rel(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


@pytest.mark.parametrize("case", [BIG[0], BIG[1], BIG[3]], ids=["sphere", "prolate", "dipoles"])
def test_group_walk_refines_the_reference(case):
    """POTG_MIXED_GROUP: each target's interactions refine the reference's
    (cells taken whole by the reference may be taken in parts), so against
    the softened float64 direct sum its errors are smaller than the
    reference walk's."""
    sim = oracle_sim(case)
    parts = sim.particles()
    ref = gpu_ctx(case, parts)
    ref.build_tree()
    grp = gpu_ctx(case, parts, precision=G.POTG_MIXED_GROUP)
    grp.build_tree()
    idx = np.arange(1, len(parts), 5, dtype=np.uint32)
    xyz = np.stack([parts["x"][idx], parts["y"][idx], parts["z"][idx]], 1)
    Ed = ref.direct_sum(0, idx)
    e_ref = _rel(ref.field_at(0, xyz, exclude=idx), Ed)
    e_grp = _rel(grp.field_at(0, xyz, exclude=idx), Ed)
    assert np.median(e_grp) < np.median(e_ref), (np.median(e_grp), np.median(e_ref))
    assert np.percentile(e_grp, 90) < np.percentile(e_ref, 90)


def test_group_walk_at_theta_zero_is_the_direct_sum():
    """At theta = 0 no cell passes the criterion: the group walk is the
    softened direct sum over all other particles, in float32."""
    case = dict(BIG[0], theta=0.0)
    parts = oracle_sim(case).particles()
    grp = gpu_ctx(case, parts, precision=G.POTG_MIXED_GROUP)
    grp.build_tree()
    idx = np.arange(1, len(parts), 11, dtype=np.uint32)
    xyz = np.stack([parts["x"][idx], parts["y"][idx], parts["z"][idx]], 1)
    e = _rel(grp.field_at(0, xyz, exclude=idx), grp.direct_sum(0, idx))
    assert np.median(e) < 1e-5 and e.max() < 1e-3, (np.median(e), e.max())


def test_group_walk_steps_and_tunes():
    case = BIG[1]
    parts = oracle_sim(case).particles()
    ctx = gpu_ctx(case, parts, precision=G.POTG_MIXED_GROUP, seed=9)
    for _ in range(3):
        ctx.step()
    t = ctx.tune()
    assert t["walk_ms"] > 0
    ctx.step()
    p, _ = ctx.download(0)
    assert np.array_equal(p["species"], parts["species"])
    assert np.all(np.isfinite(p["x"])) and ctx.stats()["steps"] == 4


@pytest.mark.parametrize("integrator,bounce", [(0, None), (1, None), (2, 0.95)],
                         ids=["boris", "verlet", "euler_richardson_bounce"])
@pytest.mark.parametrize("precision", [G.POTG_FP64, G.POTG_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("counting", [True, False], ids=["counted", "uncounted"])
def test_graph_replay_equals_eager_steps(integrator, bounce, precision, counting):
    """potg_run replays captured steps (two graphs alternating the state
    buffers); they must equal steps enqueued one by one, for every
    integrator (the two-field ones capture a second walk), both precisions,
    and with the interaction counters on or off (the batched sweep runs
    uncounted graphs)."""
    case = dict(big_case(20001, seed=21), integrator=integrator, bounce=bounce)
    parts = oracle_sim(case).particles()
    a = gpu_ctx(case, parts, precision=precision, seed=5)
    b = gpu_ctx(case, parts, precision=precision, seed=5)
    for c in (a, b):
        c.set_counting(counting)
    for _ in range(6):
        a.step()
    b.run(6)
    pa, ga = a.download(0)
    pb, gb = b.download(0)
    for f in ("x", "y", "z", "vx", "vy", "vz", "species"):
        assert np.array_equal(pa[f], pb[f]), f
    # the grain too: charge, mass, velocity (the absorbed momentum is an
    # exact integer sum, so its order does not matter), chi_b
    assert ga == gb


def test_batched_instances_equal_single_runs():
    cases = [big_case(3001, seed=31), big_case(5001, geometry=1, A=2.0, seed=32),
             big_case(2001, te=11600.0, ti=1160.0, B=(0.0, 0.0, 20.0), seed=33)]
    states = [oracle_sim(c).particles() for c in cases]
    from gpu_common import gpu_config
    batch = G.Context([gpu_config(c, seed=40 + k) for k, c in enumerate(cases)])
    for k, p in enumerate(states):
        batch.upload(k, p)
    singles = [gpu_ctx(c, states[k], seed=40 + k) for k, c in enumerate(cases)]
    for _ in range(4):
        batch.step()
        for s in singles:
            s.step()
    for k in range(len(cases)):
        pb, gb = batch.download(k)
        ps, gs = singles[k].download(0)
        for f in ("x", "y", "z", "vx", "vy", "vz"):
            assert np.array_equal(pb[f], ps[f]), (k, f)
        assert gb == gs, k


def test_many_instances_keep_one_step_count():
    """More instances than one wavefront holds (72): every instance's step
    starts from the same step count, so the ion sub-cycle (tss 42 at
    Te = Ti) falls on the same steps in each, and every instance equals its
    single run over 43 steps, across the steps (42, 43) where ions are and
    are not due.  With so few particles (601) chi_b, measured at steps 20
    and 40, can bar a species (the reference's pathology, pot.c:583-598):
    batch and singles retire such an instance (potg_set_retire), at the
    same step."""
    from gpu_common import gpu_config
    case = big_case(601, seed=81)
    cfgs = [gpu_config(case, seed=90 + k) for k in range(72)]
    state = oracle_sim(case).particles()
    batch = G.Context(cfgs, G.POTG_FP64)
    batch.set_counting(False)
    batch.set_retire(True)
    for k in range(72):
        batch.upload(k, state)
    singles = []
    for k in range(72):
        s = G.Context([cfgs[k]], G.POTG_FP64)
        s.set_counting(False)
        s.set_retire(True)
        s.upload(0, state)
        singles.append(s)
    pushed = []
    for _ in range(43):
        pushed.append([o["n_pushed"] for o in batch.step()])
        for s in singles:
            s.step()
    live = [k for k in range(72) if not batch.retired(k)[0]]
    assert len(live) >= 36, live
    k = live[-1]
    assert pushed[41][k] > pushed[40][k] and pushed[42][k] < pushed[41][k]  # ions due at 42
    for k, s in enumerate(singles):
        assert batch.retired(k) == s.retired(0), k
        pb, gb = batch.download(k)
        ps, gs = s.download(0)
        assert np.array_equal(pb.view(np.uint8), ps.view(np.uint8)), k
        assert gb == gs, k
        s.close()


def test_upload_order_and_reupload_keep_instances_apart():
    """Instances uploaded in any order, one of them re-uploaded after
    stepping: each instance's particles stay its own (the upload returns the
    buffers to reference layout before writing)."""
    from gpu_common import gpu_config
    cases = [big_case(1001, geometry=1, A=2.0, seed=51), big_case(1501, seed=52),
             big_case(801, seed=53)]
    states = [oracle_sim(c).particles() for c in cases]
    ctx = G.Context([gpu_config(c, seed=60 + k) for k, c in enumerate(cases)])
    for k in (2, 0, 1):
        ctx.upload(k, states[k])
    ctx.step()
    ctx.upload(1, states[1])
    for k in range(3):
        p, _ = ctx.download(k)
        assert len(p) == len(states[k])
        if k == 1:
            for f in ("x", "y", "z", "vx", "vy", "vz", "species"):
                assert np.array_equal(p[f], states[1][f]), f
        else:
            assert np.array_equal(p["species"], states[k]["species"])
            assert not np.array_equal(p["x"], states[k]["x"])  # stepped once


def test_bounce_off_the_wall_matches_oracle():
    """OUT_BOUNCE (pot.c:2036-2069): an electron just outside the wall and
    moving inwards is left where it is and reported (the reference warns,
    pot.c:2724-2728); one moving outwards has its velocity reflected in
    spherical components and goes back to its start position.  Both against
    the oracle."""
    case = next(c for c in CASES if c["name"] == "sphere_bounce")
    sim = oracle_sim(case, inplace=0, rng_kind=O.RNG_PHILOX, philox_seed=5)
    parts = sim.particles().copy()
    R = float(case["args"]["-r"]) * float(case["args"]["-m"])
    N = len(parts)
    for j, (fac, sign) in ((N - 1, (1.02, -1.0)), (N - 2, (1.01, 1.0))):
        assert parts["species"][j] == 2  # electrons: stepped every step
        d = np.array([parts["x"][j], parts["y"][j], parts["z"][j]]) - R
        u = d / np.linalg.norm(d)
        speed = np.sqrt(parts["vx"][j] ** 2 + parts["vy"][j] ** 2 + parts["vz"][j] ** 2)
        parts["x"][j], parts["y"][j], parts["z"][j] = R + fac * R * u
        parts["vx"][j], parts["vy"][j], parts["vz"][j] = sign * speed * u
    sim.set_particles(parts)
    ctx = gpu_ctx(case, parts, seed=5)
    oo = sim.run_step()
    og = ctx.step()[0]
    assert oo.n_wall_inward == 1 and og["n_wall_inward"] == 1
    po = sim.particles()
    pg, _ = ctx.download(0)
    # the outward electron bounced back to its start, velocity reflected
    assert po["x"][N - 2] == parts["x"][N - 2] and pg["x"][N - 2] == parts["x"][N - 2]
    v_r = (po["vx"][N - 2] * (parts["x"][N - 2] - R) + po["vy"][N - 2] * (parts["y"][N - 2] - R) +
           po["vz"][N - 2] * (parts["z"][N - 2] - R))
    assert v_r < 0
    vscale = np.abs(po["vx"]).max()
    for f in ("x", "y", "z"):
        assert np.abs(pg[f] - po[f]).max() <= 1e-12 * R, f
    for f in ("vx", "vy", "vz"):
        assert np.abs(pg[f] - po[f]).max() <= 1e-10 * vscale, f


@pytest.mark.parametrize("name", ["sphere_verlet_bz", "sphere_er_bounce_bx", "prolate_verlet"])
def test_two_field_integrators_over_steps(name):
    """Velocity Verlet and Euler-Richardson (pot.c:2532-2639) over several
    steps against the oracle with the GPU's semantics: each step takes a
    second field at the intermediate state, with the cells of the step's
    tree and the leaves moved to the intermediate positions."""
    case = next(c for c in CASES if c["name"] == name)
    sim, ctx, oo, og = _step_both(case, steps=5)
    for a, b in zip(oo, og):
        assert b["abs_q"] == a.abs_q and b["n_pushed"] == a.n_pushed
        assert b["n_wall_inward"] == a.n_wall_inward
    po = sim.particles()
    pg, gr = ctx.download(0)
    R = float(case["args"]["-r"]) * float(case["args"]["-m"])
    p0 = oracle_sim(case).particles()
    moved = np.sqrt((po["x"] - p0["x"]) ** 2 + (po["y"] - p0["y"]) ** 2 +
                    (po["z"] - p0["z"]) ** 2) > 0.05 * R
    vscale = np.abs(po["vx"]).max()
    for f in ("x", "y", "z"):
        d = np.abs(pg[f] - po[f])
        assert d[~moved].max() <= 1e-11 * R, f
    for f in ("vx", "vy", "vz"):
        d = np.abs(pg[f] - po[f])
        assert d[~moved].max() <= 1e-9 * vscale, f
    assert gr["q"] == sim.state()["q"]


@pytest.mark.parametrize("precision", [G.POTG_FP64, G.POTG_MIXED], ids=["fp64", "mixed"])
def test_replays_interleaved_with_eager_steps(precision):
    """The pot program runs the steps between its outputs as replays and the
    others one by one: any interleaving of the two ends where eager steps
    end, across a chi_b step (every 20 steps) and steps with the potential
    energies (compute_epe)."""
    case = big_case(20001, seed=23)
    parts = oracle_sim(case).particles()
    a = gpu_ctx(case, parts, precision=precision, seed=7)
    b = gpu_ctx(case, parts, precision=precision, seed=7)
    for c in (a, b):
        c.set_counting(False)
    plan = [("step", 1), ("step", 0), ("run", 4), ("step", 1), ("run", 17), ("step", 0),
            ("run", 2), ("step", 1)]
    for kind, arg in plan:
        if kind == "step":
            a.step(arg)
            b.step(arg)
        else:
            for _ in range(arg):
                a.step(0)
            b.run(arg)
    assert sum(n if k == "run" else 1 for k, n in plan) == 28
    pa, ga = a.download(0)
    pb, gb = b.download(0)
    for f in ("x", "y", "z", "vx", "vy", "vz", "species"):
        assert np.array_equal(pa[f], pb[f]), f
    assert ga == gb
