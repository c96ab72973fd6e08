"""GPU parity of the int8 matrix-core forms (and the vector-unit BaseConv)
against the CPU oracle, bit for bit, at N = 2^16.

* Forward NTT as the 4-step 256 x 256 transform on v_mfma_i32_16x16x64_i8
  (ck_plan_ntt with CK_OPT_PLAN_NTT_FORM = 1): equal to the oracle's
  Longa-Naehrig transform (NTTPhase1/NTTPhase2, src/core/NTT.cu:231-454)
  for every block shape and chunking, with and without Montgomery
  conversion, on the 24 + 8 chain and on bootparam_40's terminal/main/aux
  primes, and on the 64 polynomial x 32 limb batch (a seeded sample of 8 of
  the 64 checked, with 0, q - 1, (q - 1) / 2 and (q + 1) / 2 planted).
* BaseConv (ModSwitchMatrixMult, src/core/ModSwitch.cu:23-171) on the int8
  matrix cores and on the vector unit (ck_plan_bconv with
  CK_OPT_PLAN_BCONV_FORM = 1 / 2): 8 -> 32 limbs on 1, 6 and 64 polynomials
  (sample of 8), 12 -> 20 and 16 -> 43 limbs, with +-(s - 1) / 2 planted.
* Autotuned plans keep a candidate whose outputs are the same.
"""
import numpy as np
import pytest

from conftest import SEED_BASE
from gpu_util import assert_same, uniform

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ck():
    import cheddar_amd
    return cheddar_amd


@pytest.fixture(scope="module")
def ctx24(ck, rns24):
    return ck.Context.from_json(rns24)


@pytest.fixture(scope="module")
def ctx40(ck, boot40):
    return ck.Context.from_json(boot40)


def sync():
    torch.cuda.synchronize()


def plant_extremes(P, x, pidx):
    """coefficients 0..3 of every limb: 0, q - 1, (q - 1) / 2, (q + 1) / 2"""
    for i, pi in enumerate(pidx):
        q = P.prime(pi)
        x[i, :4] = (0, q - 1, (q - 1) // 2, (q + 1) // 2)
    return x


def check_ntt_plan(ck, ctx, P, np3, n, seed, check=None, mont=False):
    pidx = P.np_primes(*np3)
    xs = [plant_extremes(P, uniform(P, pidx, seed + k), pidx) for k in range(n)]
    d_src = [ck.to_device(x) for x in xs]
    d_dst = [torch.full_like(d, -1) for d in d_src]
    plan = ck.Plan.ntt(ctx, d_src, d_dst, np3, inverse=False, mont_conv=mont)
    plan.run()
    sync()
    for k in (range(n) if check is None else check):
        want = P.ntt_np(xs[k], *np3, mont_conv=mont)
        assert_same(ck.to_host(d_dst[k]), want, f"int8 NTT poly {k}")
    return plan


@pytest.mark.parametrize("blocks,chunk,mont", [(16, 0, False), (16, 0, True), (8, 0, False),
                                               (4, 64, True), (2, 32, False), (1, 96, False)])
def test_ntt8_plan_rns24(ck, ctx24, oracle_rns24, blocks, chunk, mont):
    ctx24.set_plan_forms(ntt="int8", ntt8_blocks=blocks, ntt8_chunk=chunk)
    try:
        plan = check_ntt_plan(ck, ctx24, oracle_rns24, (24, 0, 8), 5,
                              SEED_BASE + 3 * 1000 + blocks, mont=mont)
        info = plan.info()
        nchunk = 1 if chunk == 0 else -(-5 * 32 // chunk)
        assert info["form"] == 1 and info["launches"] == 2 * nchunk
    finally:
        ctx24.set_plan_forms(ntt="int8", ntt8_blocks=16, ntt8_chunk=0)


@pytest.mark.parametrize("np3", [(12, 0, 3), (1, 0, 0), (0, 0, 8)])
def test_ntt8_plan_partial_np(ck, ctx24, oracle_rns24, np3):
    ctx24.set_plan_forms(ntt="int8")
    try:
        check_ntt_plan(ck, ctx24, oracle_rns24, np3, 3, SEED_BASE + 3 * 1000 + 40 + sum(np3))
    finally:
        ctx24.set_plan_forms(ntt="int8")


def test_ntt8_plan_boot40_terminal_primes(ck, ctx40, oracle_boot40):
    """bootparam_40: terminal (~2^25), main (~2^30) and auxiliary primes"""
    ctx40.set_plan_forms(ntt="int8", ntt8_blocks=4)
    try:
        check_ntt_plan(ck, ctx40, oracle_boot40, (43, 4, 12), 2, SEED_BASE + 3 * 1000 + 60,
                       mont=True)
    finally:
        ctx40.set_plan_forms(ntt="int8", ntt8_blocks=16)


def test_ntt8_plan_config3_batch(ck, ctx24, oracle_rns24):
    """the 64 polynomial x 32 limb batch in one plan; a seeded sample of 8
    polynomials is checked against the oracle"""
    rng = np.random.default_rng(SEED_BASE + 3)
    sample = sorted(rng.choice(64, size=8, replace=False).tolist())
    ctx24.set_plan_forms(ntt="int8")
    try:
        check_ntt_plan(ck, ctx24, oracle_rns24, (24, 0, 8), 64, SEED_BASE + 3 * 1000 + 100,
                       check=sample)
    finally:
        ctx24.set_plan_forms(ntt="int8")


def test_ntt8_autotune(ck, ctx24, oracle_rns24):
    ctx24.set_plan_forms(ntt="auto")
    try:
        plan = check_ntt_plan(ck, ctx24, oracle_rns24, (24, 0, 8), 5,
                              SEED_BASE + 3 * 1000 + 200)
        tuning = plan.tuning()
        forms = {t[0] for t in tuning}
        assert forms == {0, 1} and all(t[3] > 0 for t in tuning)
        best = min(tuning, key=lambda t: t[3])
        assert plan.info()["form"] == best[0]
    finally:
        ctx24.set_plan_forms(ntt="int8")


def centred_sources(P, src_p, n, seed):
    rng = np.random.default_rng(seed)
    srcs = []
    for _ in range(n):
        v = np.empty((len(src_p), P.n), dtype=np.int32)
        for i, pi in enumerate(src_p):
            h = (P.prime(pi) - 1) // 2
            v[i] = rng.integers(-h, h + 1, size=P.n, dtype=np.int64)
            v[i, :3] = (h, -h, 0)
        srcs.append(v)
    return srcs


def check_bconv_plan(ck, ctx, P, src_p, dst_p, n, seed, check=None):
    _, table = P.bconv_tables(src_p, dst_p)
    srcs = centred_sources(P, src_p, n, seed)
    d_src = [torch.from_numpy(v).to("cuda") for v in srcs]
    d_dst = [torch.full((len(dst_p), P.n), -1, dtype=torch.int32, device="cuda")
             for _ in range(n)]
    plan = ck.Plan.bconv(ctx, src_p, dst_p, d_src, d_dst)
    plan.run()
    sync()
    for k in (range(n) if check is None else check):
        want = P.bconv(srcs[k], src_p, dst_p, table)
        assert_same(ck.to_host(d_dst[k]), want, f"bconv poly {k}")
    return plan


@pytest.mark.parametrize("form", ["int8", "fp64", "valu"])
@pytest.mark.parametrize("npoly", [1, 6])
def test_bconv_plan_forms_8_to_32(ck, ctx24, oracle_rns24, form, npoly):
    ctx24.set_plan_forms(bconv=form)
    try:
        plan = check_bconv_plan(ck, ctx24, oracle_rns24, list(range(24, 32)), list(range(32)),
                                npoly, SEED_BASE + 3 * 1000 + 300 + npoly)
        assert plan.info()["form"] == ck.BCONV_FORMS[form]
    finally:
        ctx24.set_plan_forms(bconv="int8")


@pytest.mark.parametrize("form", ["int8", "fp64", "valu"])
def test_bconv_plan_forms_12_to_20(ck, ctx24, oracle_rns24, form):
    """two int8 k-steps (12 sources), a partly filled second row tile"""
    ctx24.set_plan_forms(bconv=form)
    try:
        check_bconv_plan(ck, ctx24, oracle_rns24, list(range(20, 32)), list(range(20)), 2,
                         SEED_BASE + 3 * 1000 + 320)
    finally:
        ctx24.set_plan_forms(bconv="int8")


@pytest.mark.parametrize("form", ["int8", "fp64", "valu"])
def test_bconv_plan_forms_16_to_43(ck, ctx40, oracle_boot40, form):
    """16 sources, 43 destinations (three row tiles) of bootparam_40"""
    ctx40.set_plan_forms(bconv=form)
    try:
        src = list(range(0, 59, 4))[:16]
        dst = [i for i in range(59) if i not in src]
        check_bconv_plan(ck, ctx40, oracle_boot40, src, dst, 2, SEED_BASE + 3 * 1000 + 340)
    finally:
        ctx40.set_plan_forms(bconv="int8")


def test_bconv_plan_int8_config3_batch(ck, ctx24, oracle_rns24):
    """8 -> 32 limbs on the 64-polynomial batch; a seeded sample of 8 checked"""
    rng = np.random.default_rng(SEED_BASE + 3 + 1)
    sample = sorted(rng.choice(64, size=8, replace=False).tolist())
    ctx24.set_plan_forms(bconv="int8")
    try:
        check_bconv_plan(ck, ctx24, oracle_rns24, list(range(24, 32)), list(range(32)), 64,
                         SEED_BASE + 3 * 1000 + 360, check=sample)
    finally:
        ctx24.set_plan_forms(bconv="int8")


def test_bconv_autotune(ck, ctx24, oracle_rns24):
    ctx24.set_plan_forms(bconv="auto")
    try:
        plan = check_bconv_plan(ck, ctx24, oracle_rns24, list(range(24, 32)), list(range(32)),
                                4, SEED_BASE + 3 * 1000 + 380)
        tuning = plan.tuning()
        assert sorted(t[0] for t in tuning) == [0, 1, 2]
        assert plan.info()["form"] == min(tuning, key=lambda t: t[3])[0]
    finally:
        ctx24.set_plan_forms(bconv="int8")


def test_ntt_plan_in_place(ck, ctx24, oracle_rns24):
    """in place (dst = src): the int8 column pass would overwrite inputs with
    its tiled intermediate, so the plan takes the vector form; an autotuned
    in-place plan is refused"""
    P = oracle_rns24
    np3 = (24, 0, 8)
    pidx = P.np_primes(*np3)
    xs = [plant_extremes(P, uniform(P, pidx, SEED_BASE + 3 * 1000 + 400 + k), pidx)
          for k in range(3)]
    d = [ck.to_device(x) for x in xs]
    ctx24.set_plan_forms(ntt="int8")
    try:
        plan = ck.Plan.ntt(ctx24, d, d, np3, inverse=False, mont_conv=True)
        assert plan.info()["form"] == 0
        plan.run()
        sync()
        for k in range(3):
            assert_same(ck.to_host(d[k]), P.ntt_np(xs[k], *np3, mont_conv=True),
                        f"in-place NTT poly {k}")
        ctx24.set_plan_forms(ntt="auto")
        with pytest.raises(ck.CheddarError):
            ck.Plan.ntt(ctx24, d, d, np3, inverse=False)
    finally:
        ctx24.set_plan_forms(ntt="int8")
