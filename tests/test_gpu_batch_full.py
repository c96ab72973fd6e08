"""GPU batched plans at the full batched-workload shape: 64 ciphertexts with
limb counts uniform in [8, 24] (the seeded draw the benchmark uses) through
one HMult plan, then 16 hoisted rotations of every product through one plan
(autotuned BaseConv form, as benchmarked). A seeded sample of 4 items of
each plan (2 rotations per sampled item) is compared word for word with the
oracle's Context::HMult / HRot (src/core/Context.cpp:776-791, :862-871).
Also: a hoisted-rotation plan captured into a HIP graph replays to the eager
result, with eager work and synchronisation between replays."""
import random

import numpy as np
import pytest

from conftest import SEED_BASE
from gpu_util import assert_same, dev_key, full_key, uniform

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

BATCH, ROTS = 64, 16


@pytest.fixture(scope="module")
def ck():
    import cheddar_amd
    return cheddar_amd


@pytest.fixture(scope="module")
def batch(ck, rns24, oracle_rns24):
    P = oracle_rns24
    ctx = ck.Context.from_json(rns24)
    ctx.set_plan_autotune(True)
    rng = random.Random(0x4348454444415202)
    levels = [rng.randint(8, 24) - 1 for _ in range(BATCH)]
    rots = rng.sample(range(1, 1 << 15), ROTS)
    cts1 = [[uniform(P, P.level_primes(lv), SEED_BASE + 2 * 10000 + 4 * i + j)
             for j in range(2)] for i, lv in enumerate(levels)]
    cts2 = [[uniform(P, P.level_primes(lv), SEED_BASE + 2 * 10000 + 4 * i + 2 + j)
             for j in range(2)] for i, lv in enumerate(levels)]
    mkey = full_key(P, SEED_BASE + 2 * 10000 + 900)
    rkeys = [full_key(P, SEED_BASE + 2 * 10000 + 1000 + 8 * r) for r in range(ROTS)]
    return dict(ctx=ctx, levels=levels, rots=rots, cts1=cts1, cts2=cts2, mkey=mkey,
                rkeys=rkeys, sample=sorted(rng.sample(range(BATCH), 4)), rng=rng)


def test_plan_hmult_then_hoisted_full_batch(ck, batch, oracle_rns24):
    P = oracle_rns24
    b = batch
    ctx, levels, rots = b["ctx"], b["levels"], b["rots"]
    d1 = [[ck.to_device(v) for v in c] for c in b["cts1"]]
    d2 = [[ck.to_device(v) for v in c] for c in b["cts2"]]
    prod = [[torch.full((P.num_q(lv - 1), P.n), -1, dtype=torch.int32, device="cuda")
             for _ in range(2)] for lv in levels]
    p_mult = ck.Plan.hmult(ctx, levels, prod, d1, d2, dev_key(*b["mkey"]))
    assert p_mult.info()["items"] == BATCH
    p_mult.run()
    torch.cuda.synchronize()
    for i in b["sample"]:
        want = P.hmult(levels[i], b["cts1"][i], b["cts2"][i], b["mkey"], rescale=True)
        for j in range(2):
            assert_same(ck.to_host(prod[i][j]), want[j], f"hmult item {i} poly {j}")
    outs = [[[torch.full((P.num_q(lv - 1), P.n), -1, dtype=torch.int32, device="cuda")
              for _ in range(2)] for _ in rots] for lv in levels]
    p_rot = ck.Plan.hrot_hoisted(ctx, [lv - 1 for lv in levels], prod,
                                 [dev_key(*k) for k in b["rkeys"]], rots, outs)
    assert p_rot.info()["items"] == BATCH
    p_rot.run()
    torch.cuda.synchronize()
    rng = b["rng"]
    for i in b["sample"]:
        src = [ck.to_host(t) for t in prod[i]]
        for r in sorted(rng.sample(range(ROTS), 2)):
            want = P.hrot(levels[i] - 1, src, b["rkeys"][r], rots[r])
            for j in range(2):
                assert_same(ck.to_host(outs[i][r][j]), want[j],
                            f"hoisted item {i} rotation {rots[r]} poly {j}")


def test_plan_hoisted_graph_replays(ck, rns24, oracle_rns24):
    """ck_plan_run of a hoisted-rotation plan inside torch.cuda.graph: three
    replays, with eager work and synchronisation between them, each equal to
    the eager result"""
    P = oracle_rns24
    ctx = ck.Context.from_json(rns24)
    levels = [15, 9, 22]
    rots = [3, 1000, 32000, 17]
    cts = [[ck.to_device(uniform(P, P.level_primes(lv), SEED_BASE + 2 * 10000 + 2000 + 2 * i + j))
            for j in range(2)] for i, lv in enumerate(levels)]
    keys = [dev_key(*full_key(P, SEED_BASE + 2 * 10000 + 2100 + 8 * r)) for r in range(len(rots))]
    outs = [[[torch.empty((P.num_q(lv), P.n), dtype=torch.int32, device="cuda")
              for _ in range(2)] for _ in rots] for lv in levels]
    plan = ck.Plan.hrot_hoisted(ctx, levels, cts, keys, rots, outs)
    plan.run()
    torch.cuda.synchronize()
    eager = [[[t.clone() for t in pair] for pair in row] for row in outs]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan.run(stream=s)
    torch.cuda.synchronize()
    for rep in range(3):
        for row in outs:
            for pair in row:
                pair[0].zero_()
                pair[1].zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for i in range(len(levels)):
            for r in range(len(rots)):
                for k in range(2):
                    assert torch.equal(outs[i][r][k], eager[i][r][k]), (rep, i, r, k)
