"""GPU batched plans (ck_plan_*): many ciphertexts at mixed levels through
HMult and hoisted rotations, every step one launch across the batch; each
output must equal the oracle's single-ciphertext result bit for bit
(Context::HMult, src/core/Context.cpp:862-871; HoistHandler baby step,
src/extension/Hoist.cu:578-691). Also: a plan captured into a HIP graph
replays to the same results."""
import numpy as np
import pytest

from conftest import SEED_BASE
from gpu_util import assert_same, dev_key, full_key, uniform

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ck():
    import cheddar_amd
    return cheddar_amd


@pytest.fixture(scope="module",
                params=["bconv-fused", "bconv-mfma", "bconv-int8", "autotune",
                        "key-accum-pass"])
def ctx24(ck, rns24, request):
    ctx = ck.Context.from_json(rns24)
    if request.param == "autotune":
        ctx.set_plan_autotune(True)
    else:
        ctx.set_bconv_mfma("int8" if request.param == "bconv-int8"
                           else request.param == "bconv-mfma")
    # key-accum-pass: the key inner product as its own step, not fused with
    # the ModUp's forward row stages
    ctx.fused_key_rows = request.param != "key-accum-pass"
    ctx.set_fused_key_rows(ctx.fused_key_rows)
    return ctx


@pytest.fixture(scope="module")
def key24(oracle_rns24):
    return full_key(oracle_rns24, SEED_BASE + 3000)


def _empty(P, level, n=2):
    return [torch.empty((P.num_q(level), P.n), dtype=torch.int32, device="cuda")
            for _ in range(n)]


def test_plan_hmult_mixed_levels(ck, ctx24, oracle_rns24, key24):
    P = oracle_rns24
    levels = [23, 7, 15, 1, 23]
    cts1, cts2, outs = [], [], []
    for i, lv in enumerate(levels):
        pidx = P.level_primes(lv)
        cts1.append([uniform(P, pidx, SEED_BASE + 3100 + 4 * i + j) for j in range(2)])
        cts2.append([uniform(P, pidx, SEED_BASE + 3102 + 4 * i + j) for j in range(2)])
        outs.append(_empty(P, lv - 1))
    d1 = [[ck.to_device(v) for v in c] for c in cts1]
    d2 = [[ck.to_device(v) for v in c] for c in cts2]
    plan = ck.Plan.hmult(ctx24, levels, outs, d1, d2, dev_key(*key24))
    info = plan.info()
    # 9 launches (BaseConv fused) or 11 (two BaseConv steps on the matrix
    # cores); one fewer with the key product fused into the ModUp's row stages
    # (info "form": the plan's key pass)
    assert info["items"] == len(levels)
    tuning = plan.tuning()
    chunks = 1  # five small items: one chunk by default
    if not tuning:
        assert info["form"] == int(ctx24.fused_key_rows)
    else:
        # autotuned: every BaseConv form with both key passes in one chunk,
        # then the fastest of those in two chunks (five items: at least two
        # per chunk); the fastest candidate kept
        one = [t for t in tuning if t[2] == 1]
        assert sorted((t[0], t[1]) for t in one) == [(f, k) for f in range(3) for k in (0, 1)]
        best1 = min(one, key=lambda t: t[3])
        assert [t[:3] for t in tuning if t[2] != 1] == [(best1[0], best1[1], 2)]
        assert min(t[3] for t in tuning) > 0
        best = min(tuning, key=lambda t: t[3])
        assert ("fused", "fp64", "int8")[best[0]] == info["bconv_form"] and best[1] == info["form"]
        chunks = best[2]
    # 9 or 11 launches (less one with the fused key pass) per chunk
    assert info["launches"] == chunks * ((11 if info["bconv_mfma"] else 9) - info["form"])
    for rep in range(2):
        for o in outs:
            o[0].fill_(-1)
            o[1].fill_(-1)
        plan.run()
        torch.cuda.synchronize()
        for i, lv in enumerate(levels):
            want = P.hmult(lv, cts1[i], cts2[i], key24, rescale=True)
            assert_same(ck.to_host(outs[i][0]), want[0], f"plan hmult b item {i} rep {rep}")
            assert_same(ck.to_host(outs[i][1]), want[1], f"plan hmult a item {i} rep {rep}")


@pytest.mark.parametrize("moddown", [True, False])
def test_plan_hoisted_rotations(ck, ctx24, oracle_rns24, moddown):
    P = oracle_rns24
    levels = [12, 5, 20]
    rots = [1, 777, 32767, 4096, 2, 9, 10, 11, 12]  # two fused key-product steps
    keys = [full_key(P, SEED_BASE + 3300 + 16 * r) for r in range(len(rots))]
    cts, outs = [], []
    for i, lv in enumerate(levels):
        pidx = P.level_primes(lv)
        cts.append([uniform(P, pidx, SEED_BASE + 3200 + 2 * i + j) for j in range(2)])
        limbs = P.num_q(lv) + (0 if moddown else P.alpha)
        outs.append([[torch.empty((limbs, P.n), dtype=torch.int32, device="cuda")
                      for _ in range(2)] for _ in rots])
    d = [[ck.to_device(v) for v in c] for c in cts]
    plan = ck.Plan.hrot_hoisted(ctx24, levels, d, [dev_key(*k) for k in keys], rots, outs,
                                moddown=moddown)
    plan.run()
    torch.cuda.synchronize()
    for i, lv in enumerate(levels):
        for r, rot in enumerate(rots):
            if moddown:
                want = P.hrot(lv, cts[i], keys[r], rot)
            else:
                want = P.hoisted_baby_step(lv, cts[i], keys[r], rot)
            assert_same(ck.to_host(outs[i][r][0]), want[0], f"item {i} rot {rot} b")
            assert_same(ck.to_host(outs[i][r][1]), want[1], f"item {i} rot {rot} a")


def test_plan_graph_replays(ck, ctx24, oracle_rns24, key24):
    """ck_plan_run inside torch.cuda.graph: three replays, with eager work
    and synchronisation between them, each equal to the eager result."""
    P = oracle_rns24
    levels = [10, 18]
    cts1 = [[ck.to_device(uniform(P, P.level_primes(lv), SEED_BASE + 3400 + 4 * i + j))
             for j in range(2)] for i, lv in enumerate(levels)]
    cts2 = [[ck.to_device(uniform(P, P.level_primes(lv), SEED_BASE + 3402 + 4 * i + j))
             for j in range(2)] for i, lv in enumerate(levels)]
    outs = [_empty(P, lv - 1) for lv in levels]
    plan = ck.Plan.hmult(ctx24, levels, outs, cts1, cts2, dev_key(*key24))
    plan.run()
    torch.cuda.synchronize()
    eager = [[o.clone() for o in pair] for pair in outs]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan.run(stream=s)
    torch.cuda.synchronize()
    for rep in range(3):
        for pair in outs:
            pair[0].zero_()
            pair[1].zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for i in range(len(levels)):
            for k in range(2):
                assert torch.equal(outs[i][k], eager[i][k]), (rep, i, k)


def test_autotuned_plans_reject_aliasing(ck, rns24, oracle_rns24, key24):
    """the timed candidate runs of an autotuned plan write its outputs, so
    outputs that alias inputs are refused at plan creation"""
    P = oracle_rns24
    ctx = ck.Context.from_json(rns24)
    ctx.set_plan_autotune(True)
    ct = [ck.to_device(uniform(P, P.level_primes(9), SEED_BASE + 3000 + 900 + j))
          for j in range(2)]
    other = [ck.to_device(uniform(P, P.level_primes(9), SEED_BASE + 3000 + 910 + j))
             for j in range(2)]
    with pytest.raises(ck.CheddarError):
        ck.Plan.hmult(ctx, [9], [[ct[0], ct[1]]], [ct], [other], dev_key(*key24))
    with pytest.raises(ck.CheddarError):
        ck.Plan.hrot_hoisted(ctx, [9], [ct], [dev_key(*key24)], [5], [[[ct[0], other[1]]]])


@pytest.mark.parametrize("form", ["bconv-fused", "bconv-int8"])
def test_plan_hmult_terminal_primes(ck, boot40, oracle_boot40, form):
    """HMult plan on bootparam_40 at levels whose ModDownAndRescale drops
    terminal primes (28 -> 27), drops main primes and restores terminal ones
    (3 -> 2), drops terminal primes and restores main ones (2 -> 1), and a
    plain main-prime level: the restore limbs' epilogue takes the padding
    constant, and limbs with no second operand take zero; BaseConv fused into
    the forward column pass, or on the int8 matrix cores as its own step."""
    P = oracle_boot40
    ctx = ck.Context.from_json(boot40)
    ctx.set_bconv_mfma("int8" if form == "bconv-int8" else False)
    key = full_key(P, SEED_BASE + 2000)
    levels = [28, 3, 2, 9]
    cts1, cts2, outs = [], [], []
    for i, lv in enumerate(levels):
        pidx = P.level_primes(lv)
        cts1.append([uniform(P, pidx, SEED_BASE + 3500 + 4 * i + j) for j in range(2)])
        cts2.append([uniform(P, pidx, SEED_BASE + 3502 + 4 * i + j) for j in range(2)])
        outs.append(_empty(P, lv - 1))
    d1 = [[ck.to_device(v) for v in c] for c in cts1]
    d2 = [[ck.to_device(v) for v in c] for c in cts2]
    plan = ck.Plan.hmult(ctx, levels, outs, d1, d2, dev_key(*key))
    assert plan.info()["bconv_form"] == ("int8" if form == "bconv-int8" else "fused")
    plan.run()
    torch.cuda.synchronize()
    for i, lv in enumerate(levels):
        want = P.hmult(lv, cts1[i], cts2[i], key, rescale=True)
        assert_same(ck.to_host(outs[i][0]), want[0], f"boot40 plan hmult b level {lv}")
        assert_same(ck.to_host(outs[i][1]), want[1], f"boot40 plan hmult a level {lv}")
