"""GPU parity: every entry point of the HIP library (through its C ABI)
against the CPU oracle on the same seeded inputs, bit for bit, at the full
ring dimension N = 2^16.

Parameter sets: the 24 + 8 prime chain of the plain key-switching workload
(q_i = 1 mod 2^17, largest primes below 2^31) and the reference's
bootparam_40.json (4 terminal, 43 main, 12 auxiliary primes),
bootparam_35.json (5 terminal, 43 main) and bootparam_30.json (48 main).
"""
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


# every test runs with BaseConv fused into the forward NTT column pass, with
# BaseConv on the FP64 matrix cores, with BaseConv on the int8 matrix cores,
# and (BaseConv fused) with the key inner product as its own accumulation
# pass instead of fused with the ModUp's row stages: outputs must be
# identical
CONFIGS = ["bconv-fused", "bconv-mfma", "bconv-int8", "key-accum-pass"]


def configure(ctx, name):
    ctx.set_bconv_mfma("int8" if name == "bconv-int8" else name == "bconv-mfma")
    ctx.set_fused_key_rows(name != "key-accum-pass")
    return ctx


@pytest.fixture(scope="module", params=CONFIGS)
def ctx24(ck, rns24, request):
    return configure(ck.Context.from_json(rns24), request.param)


@pytest.fixture(scope="module", params=CONFIGS)
def ctx40(ck, boot40, request):
    return configure(ck.Context.from_json(boot40), request.param)


def sync():
    torch.cuda.synchronize()


# ---------------------------------------------------------------------------
# NTTHandler
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mont_conv", [False, True])
def test_ntt_forward_all_primes(ck, ctx24, oracle_rns24, mont_conv):
    P = oracle_rns24
    x = uniform(P, P.np_primes(24, 0, 8), SEED_BASE + 1)
    want = P.ntt_np(x, 24, 0, 8, mont_conv=mont_conv)
    d = ck.to_device(x)
    out = torch.empty_like(d)
    ctx24.ntt(out, d, 24, 0, 8, mont_conv=mont_conv)
    sync()
    assert_same(ck.to_host(out), want, "ntt")


def test_ntt_in_place_and_extremes(ck, ctx24, oracle_rns24):
    P = oracle_rns24
    pidx = P.np_primes(24, 0, 0)
    q = np.array([P.prime(i) for i in pidx], dtype=np.uint64)[:, None]
    for x in (np.zeros((24, P.n), np.uint32), (q - 1).astype(np.uint32) + np.zeros((24, P.n), np.uint32)):
        want = P.ntt_np(x, 24, 0, 0)
        d = ck.to_device(x)
        ctx24.ntt(d, d, 24, 0, 0)
        sync()
        assert_same(ck.to_host(d), want, "ntt in place")


def test_ntt_level_slice_with_gap(ck, ctx24, oracle_rns24):
    """NP (7 main, 0 ter, 8 aux) read from a buffer whose aux limbs sit
    after a 17-limb gap (the reference's extra_ view of a max-level key)."""
    P = oracle_rns24
    full = uniform(P, P.np_primes(24, 0, 8), SEED_BASE + 2)
    sub = np.concatenate([full[:7], full[24:]])
    want = P.ntt_np(sub, 7, 0, 8)
    d = ck.to_device(full)
    out = torch.empty((15, P.n), dtype=torch.int32, device="cuda")
    ctx24.ntt(out, d, 7, 0, 8, src_gap=17)
    sync()
    assert_same(ck.to_host(out), want, "ntt with gap")


@pytest.mark.parametrize("mont_conv", [False, True])
def test_intt_all_primes(ck, ctx24, oracle_rns24, mont_conv):
    P = oracle_rns24
    x = uniform(P, P.np_primes(24, 0, 8), SEED_BASE + 3)
    want = P.intt_np(x, 24, 0, 8, mont_conv=mont_conv)
    d = ck.to_device(x)
    out = torch.empty_like(d)
    ctx24.intt(out, d, 24, 0, 8, mont_conv=mont_conv)
    sync()
    assert_same(ck.to_host(out), want, "intt")


@pytest.mark.parametrize("normalize", [False, True])
def test_intt_mult_const(ck, ctx24, oracle_rns24, normalize):
    P = oracle_rns24
    pidx = P.np_primes(24, 0, 0)
    x = uniform(P, pidx, SEED_BASE + 4)
    rng = np.random.default_rng(4)
    consts = [int(rng.integers(0, P.prime(i))) for i in pidx]
    want = P.intt_mult_const(x, pidx, consts, normalize)
    d = ck.to_device(x)
    out = torch.empty_like(d)
    ctx24.intt_mult_const(out, d, 24, 0, 0, consts, normalize)
    sync()
    assert_same(ck.to_host(out).view(np.int32), want, "intt_mult_const")


def test_ntt_roundtrip_all_limbs(ck, ctx24, oracle_rns24):
    P = oracle_rns24
    x = uniform(P, P.np_primes(24, 0, 8), SEED_BASE + 5)
    d = ck.to_device(x)
    t = torch.empty_like(d)
    ctx24.ntt(t, d, 24, 0, 8)
    ctx24.intt(t, t, 24, 0, 8, mont_conv=False)
    sync()
    assert_same(ck.to_host(t), x, "INTT(NTT(x))")


# ---------------------------------------------------------------------------
# ModSwitchHandler
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("level", [23, 10, 7, 0])
def test_modup(ck, ctx24, oracle_rns24, level):
    P = oracle_rns24
    nq = P.num_q(level)
    x = uniform(P, P.level_primes(level), SEED_BASE + 6 + level)
    want = P.modup(x, level)
    beta, tot = want.shape[0], want.shape[1]
    d = ck.to_device(x)
    out = torch.zeros((beta * tot, P.n), dtype=torch.int32, device="cuda")
    ctx24.modup(out, d, level)
    sync()
    assert_same(ck.to_host(out).reshape(beta, tot, P.n), want, f"modup level {level}")
    assert nq + P.alpha == tot


@pytest.mark.parametrize("level,kind", [(23, 0), (23, 1), (23, 2), (5, 0), (5, 2), (1, 1), (0, 0)])
def test_moddown(ck, ctx24, oracle_rns24, level, kind):
    P = oracle_rns24
    nq = P.num_q(level)
    nsrc = nq + (0 if kind == 1 else P.alpha)
    srcs = [uniform(P, P.level_primes(level, with_aux=(kind != 1)), SEED_BASE + 40 + 2 * j)
            for j in range(2)]
    want = [P.moddown(s, level, kind) for s in srcs]
    ds = [ck.to_device(s) for s in srcs]
    outs = [torch.empty((w.shape[0], P.n), dtype=torch.int32, device="cuda") for w in want]
    ctx24.moddown(outs, ds, level, kind)
    sync()
    for j in range(2):
        assert srcs[j].shape[0] == nsrc
        assert_same(ck.to_host(outs[j]), want[j], f"moddown kind {kind} level {level} poly {j}")


# ---------------------------------------------------------------------------
# ElementWiseHandler
# ---------------------------------------------------------------------------
def test_elementwise(ck, ctx24, oracle_rns24):
    P = oracle_rns24
    pidx = P.np_primes(24, 0, 8)
    x = uniform(P, pidx, SEED_BASE + 60)
    y = uniform(P, pidx, SEED_BASE + 61)
    dx, dy = ck.to_device(x), ck.to_device(y)
    out = torch.empty_like(dx)
    np3 = (24, 0, 8)
    for name, fn, want in (("add", ctx24.add, P.add(x, y, pidx)),
                           ("sub", ctx24.sub, P.sub(x, y, pidx)),
                           ("mult", ctx24.mult, P.mult(x, y, pidx))):
        fn([out], [dx], [dy], np3)
        sync()
        assert_same(ck.to_host(out), want, name)
    ctx24.neg([out], [dx], np3)
    sync()
    assert_same(ck.to_host(out), P.neg(x, pidx), "neg")
    consts = np.array([np.uint32(P.p_prod_mont(i)) if i < 24 else np.uint32(7) for i in pidx],
                      dtype=np.uint32)
    ctx24.mult_const([out], [dx], ck.to_device(consts[None, :])[0], np3)
    sync()
    assert_same(ck.to_host(out), P.mult_const(x, consts, pidx), "mult_const")


def test_const_ops_and_imaginary_unit(ck, ctx24, oracle_rns24):
    P = oracle_rns24
    pidx = P.np_primes(24, 0, 8)
    np3 = (24, 0, 8)
    x = uniform(P, pidx, SEED_BASE + 80)
    y = uniform(P, pidx, SEED_BASE + 81)
    rng = np.random.default_rng(SEED_BASE + 82)
    c = np.array([rng.integers(0, P.prime(i)) for i in pidx], dtype=np.uint32)
    c[0], c[1] = 0, P.prime(pidx[1]) - 1
    dx, dy = ck.to_device(x), ck.to_device(y)
    dc = ck.to_device(c[None, :])[0]
    o1, o2 = torch.empty_like(dx), torch.empty_like(dx)
    for mode in range(3):
        ctx24.add_const([o1, o2], [dx, dy], dc, np3, mode=mode)
        sync()
        assert_same(ck.to_host(o1), P.add_const(x, c, pidx, mode), f"add_const mode {mode} x")
        assert_same(ck.to_host(o2), P.add_const(y, c, pidx, mode), f"add_const mode {mode} y")
    ctx24.mult_imaginary_unit([o1, o2], [dx, dy], np3)
    sync()
    assert_same(ck.to_host(o1), P.mult_imaginary_unit(x, pidx), "imaginary unit x")
    assert_same(ck.to_host(o2), P.mult_imaginary_unit(y, pidx), "imaginary unit y")


@pytest.mark.parametrize("accumulate", [False, True])
def test_paccum_caccum(ck, ctx24, oracle_rns24, accumulate):
    """sum over 20 terms (two launches of at most 16) of products with
    polynomials (PAccum / giant-step accumulation) and with per-limb
    constants (CAccum), for two polynomials at once"""
    P = oracle_rns24
    pidx = P.np_primes(24, 0, 8)
    np3 = (24, 0, 8)
    nt = 20
    xs = [[uniform(P, pidx, SEED_BASE + 100 + 2 * t + j) for j in range(2)] for t in range(nt)]
    ys = [uniform(P, pidx, SEED_BASE + 150 + t) for t in range(nt)]
    rng = np.random.default_rng(SEED_BASE + 170)
    cs = [np.array([rng.integers(0, P.prime(i)) for i in pidx], dtype=np.uint32)
          for _ in range(nt)]
    init = [uniform(P, pidx, SEED_BASE + 180 + j) for j in range(2)]
    dx = [[ck.to_device(v) for v in term] for term in xs]
    dy = [ck.to_device(v) for v in ys]
    dc = [ck.to_device(c[None, :])[0] for c in cs]
    for kind in ("paccum", "caccum"):
        outs = [ck.to_device(v) for v in init]
        if kind == "paccum":
            ctx24.paccum(outs, dx, dy, np3, accumulate=accumulate)
        else:
            ctx24.caccum(outs, dx, dc, np3, accumulate=accumulate)
        sync()
        for j in range(2):
            want = init[j] if accumulate else np.zeros_like(init[j])
            for t in range(nt):
                prod = P.mult(xs[t][j], ys[t], pidx) if kind == "paccum" else \
                    P.mult_const(xs[t][j], cs[t], pidx)
                want = P.add(want, prod, pidx)
            assert_same(ck.to_host(outs[j]), want, f"{kind} poly {j} acc {accumulate}")


@pytest.mark.parametrize("nbs,ngs", [(11, 10), (3, 2), (16, 1)])
def test_gs_accum(ck, ctx24, oracle_rns24, nbs, ngs):
    """giant-step accumulation of a hoisted transform (GSFusedKernel,
    src/extension/Hoist.cu:80-118): every giant step's sum of baby step x
    plaintext products in one pass; missing plaintexts (None) add nothing;
    more than 8 giant steps take two launches"""
    P = oracle_rns24
    pidx = P.np_primes(24, 0, 8)
    np3 = (24, 0, 8)
    xs = [[uniform(P, pidx, SEED_BASE + 400 + 2 * t + j) for j in range(2)] for t in range(nbs)]
    rng = np.random.default_rng(SEED_BASE + 470 + nbs)
    pts = [[uniform(P, pidx, SEED_BASE + 500 + g * 16 + t) if rng.random() < 0.8 else None
            for t in range(nbs)] for g in range(ngs)]
    dx = [[ck.to_device(v) for v in term] for term in xs]
    dp = [[None if v is None else ck.to_device(v) for v in row] for row in pts]
    outs = [[ck.to_device(np.full_like(xs[0][0], 0xFFFFFFFF)) for _ in range(2)]
            for _ in range(ngs)]
    ctx24.gs_accum(outs, dx, dp, np3)
    sync()
    for g in range(ngs):
        for j in range(2):
            want = np.zeros_like(xs[0][j])
            for t in range(nbs):
                if pts[g][t] is not None:
                    want = P.add(want, P.mult(xs[t][j], pts[g][t], pidx), pidx)
            assert_same(ck.to_host(outs[g][j]), want, f"giant step {g} poly {j}")
    # an output aliasing an operand is rejected before anything is launched
    with pytest.raises(ck.CheddarError):
        ctx24.gs_accum([[dx[0][0], outs[0][1]]] + outs[1:], dx, dp, np3)


@pytest.mark.parametrize("npoly", [1, 6])
def test_bconv_plan_8_to_32(ck, ctx24, oracle_rns24, npoly):
    """standalone BaseConv plan (ModSwitchMatrixMult, src/core/ModSwitch.cu:
    23-171) from the 8 special primes to all 32 primes (the 8 -> 32 limb
    contraction) over several polynomials in one launch: equal to the
    oracle's table-driven conversion for centred sources, including the
    extreme values +-(s - 1) / 2"""
    P = oracle_rns24
    src_p = list(range(24, 32))
    dst_p = list(range(32))
    _, table = P.bconv_tables(src_p, dst_p)
    rng = np.random.default_rng(SEED_BASE + 700 + npoly)
    srcs = []
    for k in range(npoly):
        v = np.empty((len(src_p), P.n), dtype=np.int32)
        for i, pi in enumerate(src_p):
            h = (P.prime(pi) - 1) // 2
            v[i] = rng.integers(-h, h + 1, size=P.n, dtype=np.int64)
            v[i, :2] = (h, -h)
        srcs.append(v)
    d_src = [torch.from_numpy(v).to("cuda") for v in srcs]
    d_dst = [torch.full((len(dst_p), P.n), -1, dtype=torch.int32, device="cuda")
             for _ in range(npoly)]
    plan = ck.Plan.bconv(ctx24, src_p, dst_p, d_src, d_dst)
    plan.run()
    sync()
    for k in range(npoly):
        want = P.bconv(srcs[k], src_p, dst_p, table)
        assert_same(ck.to_host(d_dst[k]), want, f"bconv poly {k}")


@pytest.mark.parametrize("inverse", [False, True])
def test_ntt_plan_batch(ck, ctx24, oracle_rns24, inverse):
    """batched NTT / INTT plan (one launch per pass over 5 polynomials of all
    32 limbs) equal to the oracle's transform of each"""
    P = oracle_rns24
    pidx = P.np_primes(24, 0, 8)
    xs = [uniform(P, pidx, SEED_BASE + 750 + k) for k in range(5)]
    d_src = [ck.to_device(x) for x in xs]
    d_dst = [torch.empty_like(d) for d in d_src]
    plan = ck.Plan.ntt(ctx24, d_src, d_dst, (24, 0, 8), inverse=inverse, mont_conv=inverse)
    info = plan.info()
    assert info["items"] == 5 and info["launches"] == 2
    plan.run()
    sync()
    for k, x in enumerate(xs):
        want = (P.intt_np(x, 24, 0, 8, mont_conv=True) if inverse
                else P.ntt_np(x, 24, 0, 8, mont_conv=False))
        assert_same(ck.to_host(d_dst[k]), want, f"plan transform {k}")


@pytest.mark.parametrize("num_aux", [0, 12])
def test_boot40_modup_to_max(ck, ctx40, oracle_boot40, num_aux):
    P = oracle_boot40
    base = P.np_primes(0, 2)
    src = uniform(P, base, SEED_BASE + 90)
    dst = P.np_primes(*P.level_np(P.max_level), num_aux)
    want = P.modup_to_max(src, base, dst)
    out = torch.empty((len(dst), P.n), dtype=torch.int32, device="cuda")
    ctx40.modup_to_max(out, ck.to_device(src), (0, 2), num_aux)
    sync()
    assert_same(ck.to_host(out), want, f"modup_to_max aux {num_aux}")


def test_boot40_crt_lift(ck, ctx40, oracle_boot40):
    """exact lift between explicit primes both ways (the two-prime key
    switches of sparse-secret encapsulation): short base -> first two
    auxiliary primes, and back, and one prime -> many"""
    P = oracle_boot40
    short = P.np_primes(0, 2)
    aux = [P.num_primes - P.alpha, P.num_primes - P.alpha + 1]
    for src_p, dst_p, seed in ((short, aux, 92), (aux, short, 93), ([aux[0]], list(range(20)), 94)):
        src = uniform(P, src_p, SEED_BASE + seed)
        want = P.modup_to_max(src, src_p, dst_p)
        out = torch.empty((len(dst_p), P.n), dtype=torch.int32, device="cuda")
        ctx40.crt_lift(out, ck.to_device(src), src_p, dst_p)
        sync()
        assert_same(ck.to_host(out), want, f"crt_lift {src_p} -> {dst_p[:4]}")


def test_modup_to_max_single_prime(ck, ctx24, oracle_rns24):
    P = oracle_rns24
    base = P.np_primes(1, 0)
    src = uniform(P, base, SEED_BASE + 91)
    dst = P.np_primes(24, 0, 8)
    out = torch.empty((len(dst), P.n), dtype=torch.int32, device="cuda")
    ctx24.modup_to_max(out, ck.to_device(src), (1, 0), 8)
    sync()
    assert_same(ck.to_host(out), P.modup_to_max(src, base, dst), "modup_to_max one prime")


def test_tensor(ck, ctx24, oracle_rns24):
    P = oracle_rns24
    pidx = P.level_primes(23)
    ins = [uniform(P, pidx, SEED_BASE + 70 + i) for i in range(4)]
    want = P.tensor(*ins, pidx)
    d = [ck.to_device(v) for v in ins]
    outs = [torch.empty_like(d[0]) for _ in range(3)]
    ctx24.tensor(*outs, *d, 24, 0)
    sync()
    for i in range(3):
        assert_same(ck.to_host(outs[i]), want[i], f"tensor d{i}")


def test_tensor_square(ck, ctx24, oracle_rns24):
    """TensorSquare (ElementWise.cu:549-575): the same ciphertext as both
    operands (the squaring kernel) equals the oracle's general Tensor of the
    ciphertext with a copy of itself, including 0 and q - 1 coefficients"""
    P = oracle_rns24
    pidx = P.level_primes(23)
    b, a = (uniform(P, pidx, SEED_BASE + 75 + i) for i in range(2))
    q = np.array([P.prime(i) for i in pidx], dtype=np.uint32)
    b[:, :2] = np.stack([np.zeros_like(q), q - 1], axis=1)
    a[:, :2] = np.stack([q - 1, q - 1], axis=1)
    want = P.tensor(b, a, b.copy(), a.copy(), pidx)
    db, da = ck.to_device(b), ck.to_device(a)
    outs = [torch.empty_like(db) for _ in range(3)]
    ctx24.tensor(*outs, db, da, db, da, 24, 0)
    sync()
    for i in range(3):
        assert_same(ck.to_host(outs[i]), want[i], f"tensor square d{i}")


@pytest.mark.parametrize("rot", [1, 1234, 32767, -1])
def test_permute(ck, ctx24, oracle_rns24, rot):
    P = oracle_rns24
    pidx = P.level_primes(23)
    x = uniform(P, pidx, SEED_BASE + 80)
    want = P.permute(x, P.galois(rot))
    d = ck.to_device(x)
    out = torch.empty_like(d)
    ctx24.permute([out], [d], rot, (24, 0, 0))
    sync()
    assert_same(ck.to_host(out), want, f"permute {rot}")


def test_copy(ck, ctx24):
    """ck_copy (DeviceVector's device copies, src/core/DeviceVector.cpp:
    105-124): the 16-byte vector kernel for aligned ranges whose size is a
    multiple of 16 (one limb, three limbs and 16 bytes, 64 MiB), the
    runtime's copy for the others (12 bytes; an offset of 4 bytes; 1000
    bytes); bytes outside the range stay untouched, dst == src is a no-op
    and overlapping ranges are refused"""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(SEED_BASE + 90)
    for nbytes, off in [(1 << 18, 0), (3 << 18 | 16, 0), (64 << 20, 0), (12, 0),
                        (1 << 18, 4), (1000, 4)]:
        total = nbytes + off + 64
        src = torch.randint(0, 256, (total,), generator=gen, device="cuda",
                            dtype=torch.int32).to(torch.uint8)
        dst = torch.zeros(total, dtype=torch.uint8, device="cuda")
        ctx24.copy(dst[off:], src[off:], nbytes)
        sync()
        got, want = dst.cpu(), src.cpu()
        assert torch.equal(got[off:off + nbytes], want[off:off + nbytes]), (nbytes, off)
        assert int(got[:off].sum()) == 0 and int(got[off + nbytes:].sum()) == 0, (nbytes, off)
    buf = torch.arange(4096, device="cuda", dtype=torch.int32).to(torch.uint8)
    ctx24.copy(buf, buf, 4096)
    sync()
    assert torch.equal(buf.cpu(), torch.arange(4096, dtype=torch.int32).to(torch.uint8))
    with pytest.raises(ck.CheddarError):
        ctx24.copy(buf[16:], buf[:2048], 2048)
    with pytest.raises(ValueError):  # more bytes than the destination holds
        ctx24.copy(buf[:1024], buf[2048:], 2048)


@pytest.mark.parametrize("accumulate", [False, True])
def test_permute_accum(ck, ctx24, oracle_rns24, accumulate):
    """sum of automorphisms over Q*P limbs (PermuteAccum, src/core/
    ElementWise.cu:394-501): 10 sources (two launches of at most 8) with
    rotations, conjugation and the identity; a missing second polynomial
    adds nothing to it"""
    P = oracle_rns24
    pidx = P.np_primes(24, 0, 8)
    np3 = (24, 0, 8)
    rots = [1, 1234, 0, -1, 32767, 5, 0, 77, 4096, 3]
    srcs = [[uniform(P, pidx, SEED_BASE + 600 + 2 * s + j) for j in range(2)]
            for s in range(len(rots))]
    srcs[2][1] = None
    srcs[7][1] = None
    init = [uniform(P, pidx, SEED_BASE + 640 + j) for j in range(2)]
    d = [[None if v is None else ck.to_device(v) for v in row] for row in srcs]
    outs = [ck.to_device(v) for v in init]
    ctx24.permute_accum(outs, d, rots, np3, accumulate=accumulate)
    sync()
    for j in range(2):
        want = init[j] if accumulate else np.zeros_like(init[j])
        for s, r in enumerate(rots):
            if srcs[s][j] is not None:
                t = srcs[s][j] if r == 0 else P.permute(srcs[s][j], P.galois(r))
                want = P.add(want, t, pidx)
        assert_same(ck.to_host(outs[j]), want, f"permute_accum poly {j}")


@pytest.mark.parametrize("accumulate", [False, True])
def test_permute_accum_no_sources(ck, ctx24, oracle_rns24, accumulate):
    """an empty sum of automorphisms over all 32 limbs: zeros, or dst
    unchanged when accumulating (the masked terms gather in bounds)"""
    P = oracle_rns24
    pidx = P.np_primes(24, 0, 8)
    init = [uniform(P, pidx, SEED_BASE + 660 + j) for j in range(2)]
    outs = [ck.to_device(v) for v in init]
    ctx24.permute_accum(outs, [], [], (24, 0, 8), accumulate=accumulate)
    sync()
    for j in range(2):
        want = init[j] if accumulate else np.zeros_like(init[j])
        assert_same(ck.to_host(outs[j]), want, f"empty permute_accum poly {j}")


# ---------------------------------------------------------------------------
# Context evaluator
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def key24(oracle_rns24):
    return full_key(oracle_rns24, SEED_BASE + 1000)


@pytest.mark.parametrize("level", [23, 12])
def test_keyswitch_ops_in_place(ck, ctx24, oracle_rns24, key24, level):
    """HMult with its output over ct1's buffers (every read of ct1 precedes
    the ModDown's first write), MultKey (Context.cpp:584-607: ModDown of the
    key-switched a plus P b) and HRot, against the oracle"""
    P = oracle_rns24
    pidx = P.level_primes(level)
    ct1 = [uniform(P, pidx, SEED_BASE + 140 + i) for i in range(2)]
    ct2 = [uniform(P, pidx, SEED_BASE + 150 + i) for i in range(2)]
    dk = dev_key(*key24)
    d1 = [ck.to_device(v) for v in ct1]
    d2 = [ck.to_device(v) for v in ct2]
    nq, nq_out = P.num_q(level), P.num_q(level - 1)
    want_h = P.hmult(level, ct1, ct2, key24, rescale=True)
    acc = P.mult_key_no_moddown(level, ct1[1], key24[0], key24[1], ct1[0], None)
    want_k = [P.moddown(acc[j], level, 0) for j in range(2)]
    want_r = P.hrot(level, ct1, key24, 77)
    inp = [d.clone() for d in d1]
    ctx24.hmult([inp[0][:nq_out], inp[1][:nq_out]], inp, d2, dk, level, rescale=True)
    outk = [torch.empty((nq, P.n), dtype=torch.int32, device="cuda") for _ in range(2)]
    ctx24.mult_key(outk, d1, dk, level)
    outr = [torch.empty((nq, P.n), dtype=torch.int32, device="cuda") for _ in range(2)]
    ctx24.hrot(outr, d1, dk, 77, level)
    sync()
    for j in range(2):
        assert_same(ck.to_host(inp[j][:nq_out]), want_h[j], f"in-place hmult level {level}")
        assert_same(ck.to_host(outk[j]), want_k[j], f"mult_key level {level}")
        assert_same(ck.to_host(outr[j]), want_r[j], f"hrot level {level}")


def test_hmult_graph_replays(ck, ctx24, oracle_rns24, key24):
    """one HMult captured into a HIP graph on a caller stream, replayed three
    times with eager work and synchronisation in between: every replay
    equals the oracle"""
    P = oracle_rns24
    level = 23
    pidx = P.level_primes(level)
    ct1 = [uniform(P, pidx, SEED_BASE + 160 + i) for i in range(2)]
    ct2 = [uniform(P, pidx, SEED_BASE + 170 + i) for i in range(2)]
    want = P.hmult(level, ct1, ct2, key24, rescale=True)
    dk = dev_key(*key24)
    d1 = [ck.to_device(v) for v in ct1]
    d2 = [ck.to_device(v) for v in ct2]
    out = [torch.empty((want[0].shape[0], P.n), dtype=torch.int32, device="cuda")
           for _ in range(2)]
    st = torch.cuda.Stream()
    ctx24.hmult(out, d1, d2, dk, level, stream=st)  # warm-up outside capture
    sync()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=st):
        ctx24.hmult(out, d1, d2, dk, level, stream=st)
    for rep in range(3):
        for o in out:
            o.fill_(-1)
        sync()
        g.replay()
        sync()
        for j in range(2):
            assert_same(ck.to_host(out[j]), want[j], f"graph replay {rep} poly {j}")


@pytest.mark.parametrize("level,rescale", [(23, True), (23, False), (8, True), (1, True)])
def test_hmult(ck, ctx24, oracle_rns24, key24, level, rescale):
    P = oracle_rns24
    pidx = P.level_primes(level)
    ct1 = [uniform(P, pidx, SEED_BASE + 100 + i) for i in range(2)]
    ct2 = [uniform(P, pidx, SEED_BASE + 110 + i) for i in range(2)]
    want = P.hmult(level, ct1, ct2, key24, rescale=rescale)
    nq_out = want[0].shape[0]
    dk = dev_key(*key24)
    d1 = [ck.to_device(v) for v in ct1]
    d2 = [ck.to_device(v) for v in ct2]
    out = [torch.empty((nq_out, P.n), dtype=torch.int32, device="cuda") for _ in range(2)]
    ctx24.hmult(out, d1, d2, dk, level, rescale=rescale)
    sync()
    assert_same(ck.to_host(out[0]), want[0], f"hmult b level {level}")
    assert_same(ck.to_host(out[1]), want[1], f"hmult a level {level}")


def test_mult_key_no_moddown(ck, ctx24, oracle_rns24, key24):
    P = oracle_rns24
    level = 23
    pidx = P.level_primes(level)
    x = uniform(P, pidx, SEED_BASE + 120)
    ab = uniform(P, pidx, SEED_BASE + 121)
    aa = uniform(P, pidx, SEED_BASE + 122)
    want = P.mult_key_no_moddown(level, x, key24[0], key24[1], ab, aa)
    tot = P.num_q(level) + P.alpha
    acc = [torch.empty((tot, P.n), dtype=torch.int32, device="cuda") for _ in range(2)]
    ctx24.mult_key_no_moddown(acc[0], acc[1], ck.to_device(x), dev_key(*key24), level,
                              ck.to_device(ab), ck.to_device(aa))
    sync()
    assert_same(ck.to_host(acc[0]), want[0], "accum b")
    assert_same(ck.to_host(acc[1]), want[1], "accum a")


@pytest.mark.parametrize("level,rot", [(23, 1), (23, 1234), (12, 32767), (3, 5), (23, 0),
                                       (9, 32768)])
def test_hrot(ck, ctx24, oracle_rns24, key24, level, rot):
    P = oracle_rns24
    pidx = P.level_primes(level)
    ct = [uniform(P, pidx, SEED_BASE + 130 + i) for i in range(2)]
    want = P.hrot(level, ct, key24, rot)
    d = [ck.to_device(v) for v in ct]
    out = [torch.empty_like(d[0]) for _ in range(2)]
    ctx24.hrot(out, d, dev_key(*key24), rot, level)
    sync()
    assert_same(ck.to_host(out[0]), want[0], "hrot b")
    assert_same(ck.to_host(out[1]), want[1], "hrot a")


# ---------------------------------------------------------------------------
# bootparam_40.json: terminal primes, uneven level steps, dnum 4
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("level", [28, 16, 2, 0])
def test_boot40_modup(ck, ctx40, oracle_boot40, level):
    P = oracle_boot40
    x = uniform(P, P.level_primes(level), SEED_BASE + 200 + level)
    want = P.modup(x, level)
    beta, tot = want.shape[0], want.shape[1]
    d = ck.to_device(x)
    out = torch.zeros((beta * tot, P.n), dtype=torch.int32, device="cuda")
    ctx40.modup(out, d, level)
    sync()
    assert_same(ck.to_host(out).reshape(beta, tot, P.n), want, f"boot40 modup {level}")


# level transitions of bootparam_40: 28->27 drops terminal primes,
# 3->2 (4,2)->(1,4) drops main primes and restores terminal ones,
# 2->1 (1,4)->(3,0) drops terminal primes and restores main ones
@pytest.mark.parametrize("level,kind", [(28, 0), (28, 1), (28, 2), (27, 2), (17, 2),
                                        (3, 1), (3, 2), (2, 1), (2, 2), (1, 2), (5, 0)])
def test_boot40_moddown(ck, ctx40, oracle_boot40, level, kind):
    P = oracle_boot40
    srcs = [uniform(P, P.level_primes(level, with_aux=(kind != 1)), SEED_BASE + 300 + 2 * j)
            for j in range(2)]
    want = [P.moddown(s, level, kind) for s in srcs]
    ds = [ck.to_device(s) for s in srcs]
    outs = [torch.empty((w.shape[0], P.n), dtype=torch.int32, device="cuda") for w in want]
    ctx40.moddown(outs, ds, level, kind)
    sync()
    for j in range(2):
        assert_same(ck.to_host(outs[j]), want[j], f"boot40 moddown kind {kind} level {level}")


@pytest.mark.parametrize("level", [28, 9])
def test_boot40_hmult(ck, ctx40, oracle_boot40, level):
    P = oracle_boot40
    key = full_key(P, SEED_BASE + 2000)
    pidx = P.level_primes(level)
    ct1 = [uniform(P, pidx, SEED_BASE + 400 + i) for i in range(2)]
    ct2 = [uniform(P, pidx, SEED_BASE + 410 + i) for i in range(2)]
    want = P.hmult(level, ct1, ct2, key, rescale=True)
    d1 = [ck.to_device(v) for v in ct1]
    d2 = [ck.to_device(v) for v in ct2]
    out = [torch.empty((want[0].shape[0], P.n), dtype=torch.int32, device="cuda") for _ in range(2)]
    ctx40.hmult(out, d1, d2, dev_key(*key), level, rescale=True)
    sync()
    assert_same(ck.to_host(out[0]), want[0], "boot40 hmult b")
    assert_same(ck.to_host(out[1]), want[1], "boot40 hmult a")


# ---------------------------------------------------------------------------
# Hoisted rotations (HoistHandler baby step, src/extension/Hoist.cu)
# ---------------------------------------------------------------------------
def _rot_keys(P, n, seed):
    return [full_key(P, seed + 16 * i) for i in range(n)]


@pytest.mark.parametrize("level", [23, 7])
def test_hoisted_baby_step(ck, ctx24, oracle_rns24, level):
    P = oracle_rns24
    pidx = P.level_primes(level)
    ct = [uniform(P, pidx, SEED_BASE + 140 + i) for i in range(2)]
    rots = [1, 3, 1234, 32767]
    keys = _rot_keys(P, len(rots), SEED_BASE + 150)
    tot = P.num_q(level) + P.alpha
    outs = [[torch.empty((tot, P.n), dtype=torch.int32, device="cuda") for _ in range(2)]
            for _ in rots]
    ctx24.hrot_hoisted(outs, [ck.to_device(v) for v in ct], [dev_key(*k) for k in keys],
                       rots, level, moddown=False)
    sync()
    for r, k, o in zip(rots, keys, outs):
        want = P.hoisted_baby_step(level, ct, k, r)
        assert_same(ck.to_host(o[0]), want[0], f"bs b rot {r}")
        assert_same(ck.to_host(o[1]), want[1], f"bs a rot {r}")


def test_hoisted_rotations_equal_hrot(ck, ctx24, oracle_rns24):
    """10 rotations (two launches of the fused kernel) + ModDown == HRot"""
    P = oracle_rns24
    level = 15
    pidx = P.level_primes(level)
    ct = [uniform(P, pidx, SEED_BASE + 160 + i) for i in range(2)]
    rots = [1, 2, 5, 77, 1000, 4096, 8191, 16384, 30000, 32767]
    keys = _rot_keys(P, len(rots), SEED_BASE + 170)
    outs = [[torch.empty((P.num_q(level), P.n), dtype=torch.int32, device="cuda")
             for _ in range(2)] for _ in rots]
    ctx24.hrot_hoisted(outs, [ck.to_device(v) for v in ct], [dev_key(*k) for k in keys],
                       rots, level, moddown=True)
    sync()
    for r, k, o in zip(rots, keys, outs):
        want = P.hrot(level, ct, k, r)
        assert_same(ck.to_host(o[0]), want[0], f"hoisted hrot b rot {r}")
        assert_same(ck.to_host(o[1]), want[1], f"hoisted hrot a rot {r}")


def test_boot40_hoisted_rotations(ck, ctx40, oracle_boot40):
    """terminal primes and dnum 4: hoisted HRot == HRot"""
    P = oracle_boot40
    level = 20
    pidx = P.level_primes(level)
    ct = [uniform(P, pidx, SEED_BASE + 180 + i) for i in range(2)]
    rots = [3, 12345]
    keys = _rot_keys(P, len(rots), SEED_BASE + 190)
    outs = [[torch.empty((P.num_q(level), P.n), dtype=torch.int32, device="cuda")
             for _ in range(2)] for _ in rots]
    ctx40.hrot_hoisted(outs, [ck.to_device(v) for v in ct], [dev_key(*k) for k in keys],
                       rots, level, moddown=True)
    sync()
    for r, k, o in zip(rots, keys, outs):
        want = P.hrot(level, ct, k, r)
        assert_same(ck.to_host(o[0]), want[0], f"boot40 hoisted b rot {r}")
        assert_same(ck.to_host(o[1]), want[1], f"boot40 hoisted a rot {r}")


# ---------------------------------------------------------------------------
# bootparam_30.json (48 main primes, no terminal primes) and bootparam_35.json
# (5 terminal primes: rescales drop up to 5 terminal primes and restore up to
# 3 main ones, BaseConv from 17 sources); the reference's BasicTest runs on
# both (unittest/BasicTest.cpp:421-429)
# ---------------------------------------------------------------------------
OTHER_PARAMS = {
    "bootparam_30.json": {"modup": [34, 12, 0], "hmult": [34, 10],
                          "moddown": [(34, 0), (34, 2), (23, 2), (5, 1), (1, 2)]},
    "bootparam_35.json": {"modup": [31, 9, 3], "hmult": [31, 9],
                          "moddown": [(31, 0), (31, 2), (9, 1), (9, 2), (3, 2), (2, 1),
                                      (2, 2), (8, 2), (1, 2)]},
}


@pytest.fixture(scope="module", params=[(p, c) for p in OTHER_PARAMS
                                        for c in ("bconv-fused", "bconv-int8")],
                ids=lambda v: f"{v[0][:-5]}-{v[1]}")
def other(ck, request):
    import os
    from conftest import load_params
    import ckks_oracle as O
    name, cfg = request.param
    prm = load_params(name)
    O.set_threads(min(8, os.cpu_count() or 1))
    return name, configure(ck.Context.from_json(prm), cfg), O.Param.from_json(prm)


def test_other_params_modup(ck, other):
    name, ctx, P = other
    for level in OTHER_PARAMS[name]["modup"]:
        x = uniform(P, P.level_primes(level), SEED_BASE + 250 + level)
        want = P.modup(x, level)
        beta, tot = want.shape[0], want.shape[1]
        out = torch.zeros((beta * tot, P.n), dtype=torch.int32, device="cuda")
        ctx.modup(out, ck.to_device(x), level)
        sync()
        assert_same(ck.to_host(out).reshape(beta, tot, P.n), want, f"{name} modup {level}")


def test_other_params_moddown(ck, other):
    name, ctx, P = other
    for level, kind in OTHER_PARAMS[name]["moddown"]:
        srcs = [uniform(P, P.level_primes(level, with_aux=(kind != 1)),
                        SEED_BASE + 350 + 2 * j + level) for j in range(2)]
        want = [P.moddown(s, level, kind) for s in srcs]
        outs = [torch.empty((w.shape[0], P.n), dtype=torch.int32, device="cuda") for w in want]
        ctx.moddown(outs, [ck.to_device(s) for s in srcs], level, kind)
        sync()
        for j in range(2):
            assert_same(ck.to_host(outs[j]), want[j], f"{name} moddown kind {kind} level {level}")


def test_other_params_hmult(ck, other):
    name, ctx, P = other
    key = full_key(P, SEED_BASE + 2500)
    dk = dev_key(*key)
    for level in OTHER_PARAMS[name]["hmult"]:
        pidx = P.level_primes(level)
        ct1 = [uniform(P, pidx, SEED_BASE + 450 + i + level) for i in range(2)]
        ct2 = [uniform(P, pidx, SEED_BASE + 460 + i + level) for i in range(2)]
        want = P.hmult(level, ct1, ct2, key, rescale=True)
        out = [torch.empty((want[0].shape[0], P.n), dtype=torch.int32, device="cuda")
               for _ in range(2)]
        ctx.hmult(out, [ck.to_device(v) for v in ct1], [ck.to_device(v) for v in ct2], dk,
                  level, rescale=True)
        sync()
        for j in range(2):
            assert_same(ck.to_host(out[j]), want[j], f"{name} hmult level {level} poly {j}")
