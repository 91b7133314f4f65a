"""The float32 walk's two visit orders (potg_tuning.variant with 64 targets
per wavefront): 0 visits a range's cells until some lanes open one, 1 sweeps
each range whole before descending (k_walk32s).  Both apply each target's
own MAC to the same float32 operands, so each target meets exactly the same
cells and leaves in either order: the per-target interaction counts are
equal, and only the order of the float32 sums differs.  Against the float64
oracle the sweep keeps the mixed walk's stated tolerance (median 1e-5, 99th
percentile 2e-3: MAC decisions that flip at the threshold in float32)."""
import numpy as np
import pytest

from gpu_common import CASES, big_case, gpu_config, gpu_ctx, oracle_sim

from pot_amd import potg as G

pytestmark = pytest.mark.gpu

BIG = [big_case(20001), big_case(20001, geometry=1, A=2.0, seed=12),
       big_case(20001, geometry=2, A=0.5, seed=13),
       dict(big_case(20001, seed=14), name="big20001_dipoles", dipoles=1),
       dict(big_case(60001, seed=15, r_over_ld=12.0), name="big60001_dipoles_t015", dipoles=1,
            theta=0.15)]
ALL = list(CASES) + BIG
IDS = [c["name"] for c in ALL]


def _fields(ctx, parts, idx, variant):
    ctx.set_tuning(2048, 64, variant)
    xyz = np.stack([parts["x"][idx], parts["y"][idx], parts["z"][idx]], 1)
    return ctx.field_at(0, xyz, exclude=idx, counts=True), xyz


@pytest.mark.parametrize("case", ALL, ids=IDS)
def test_sweep_meets_the_same_interactions(case):
    sim = oracle_sim(case)
    parts = sim.particles()
    ctx = gpu_ctx(case, parts, precision=G.POTG_MIXED)
    ctx.build_tree()
    idx = np.arange(1, len(parts), max(1, len(parts) // 4000), dtype=np.uint32)
    (E0, c0), xyz = _fields(ctx, parts, idx, 0)
    (E1, c1), _ = _fields(ctx, parts, idx, 1)
    assert np.array_equal(c0, c1)
    rel01 = np.linalg.norm(E1 - E0, axis=1) / np.linalg.norm(E0, axis=1)
    assert np.median(rel01) < 2e-6 and rel01.max() < 1e-3, (np.median(rel01), rel01.max())
    sim.build_tree()
    Eo = np.array([sim.field_at(int(i), *xyz[k], 1) for k, i in enumerate(idx)])
    rel = np.linalg.norm(E1 - Eo, axis=1) / np.linalg.norm(Eo, axis=1)
    assert np.median(rel) < 1e-5, np.median(rel)
    assert np.percentile(rel, 99) < 2e-3, np.percentile(rel, 99)


@pytest.mark.parametrize("dipoles", [0, 1], ids=["monopole", "dipoles"])
def test_sweep_steps_conserve_and_count(dipoles):
    """Steps with the sweep: counting on and off end in the same state, and
    the step's interaction counts equal those of the other order."""
    case = dict(big_case(20001, seed=31), dipoles=dipoles)
    parts = oracle_sim(case).particles()
    ctxs = [gpu_ctx(case, parts, precision=G.POTG_MIXED, seed=5) for _ in range(3)]
    for c, var in zip(ctxs, (1, 1, 0)):
        c.set_tuning(2048, 64, var)
    ctxs[1].set_counting(False)
    outs = [[c.step()[0] for _ in range(3)] for c in ctxs]
    assert outs[0] == outs[1]
    pa, ga = ctxs[0].download(0)
    pb, gb = ctxs[1].download(0)
    assert np.array_equal(pa.view(np.uint8), pb.view(np.uint8)) and ga == gb
    sa, sc = ctxs[0].stats(), ctxs[2].stats()
    # the first step's tree and targets are the same in both orders; its
    # counts too (later steps part ways by float32 rounding of the pushes)
    assert sa["n_cell"] > 0 and sa["n_open"] > 0
    one = [gpu_ctx(case, parts, precision=G.POTG_MIXED, seed=5) for _ in range(2)]
    for c, var in zip(one, (0, 1)):
        c.set_tuning(2048, 64, var)
        c.step()
    s0, s1 = one[0].stats(), one[1].stats()
    for k in ("n_cell", "n_leaf", "n_open"):
        assert s0[k] == s1[k], k
    assert sc["steps"] == sa["steps"] == 3


def test_untuned_shape_follows_size():
    """Until potg_tune or potg_set_tuning the walk's shape follows the
    context's 64-target tiles per CU (half the particles are targets): from
    16 per CU the float32 walk sweeps whole ranges with 64 targets per
    wavefront on 8 blocks per CU, from 1 per CU the same walk on 2 blocks
    per CU, below that 8 targets per wavefront on 2 blocks per CU; the
    float64 walk takes 64 targets on 8 blocks per CU from 16 tiles per CU,
    4 below."""
    shapes = {}
    for name, n in (("tiny", 20001), ("middle", 200001), ("large", 1500001)):
        cfg = gpu_config(big_case(n))
        for prec in (G.POTG_MIXED, G.POTG_FP64):
            ctx = G.Context([cfg], prec)
            shapes[name, prec] = ctx.tuning()
            ctx.close()
    cus = shapes["large", G.POTG_MIXED]["walk_blocks"] // 8
    assert cus * 8 == shapes["large", G.POTG_MIXED]["walk_blocks"]
    # 200001 particles: at least one tile per CU on a device of up to 781 CUs
    want = {("large", G.POTG_MIXED): (64, 1, 8), ("middle", G.POTG_MIXED): (64, 1, 2),
            ("tiny", G.POTG_MIXED): (8, 0, 2), ("large", G.POTG_FP64): (64, 0, 8),
            ("middle", G.POTG_FP64): (64, 0, 4), ("tiny", G.POTG_FP64): (64, 0, 4)}
    for key, (group, variant, bpc) in want.items():
        t = shapes[key]
        assert (t["group"], t["variant"], t["walk_blocks"]) == (group, variant, bpc * cus), (key, t)
