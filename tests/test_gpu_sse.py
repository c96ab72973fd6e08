"""GPU parity of the level -1 key switch: sparse-secret encapsulation's
dense -> sparse switch, Context::MultKey with the DtS ModSwitchHandler
(src/core/Context.cpp:567-607, src/core/ModSwitch.cu:175-232, P' from
Context.cpp:109-123), word for word against the oracle's restatement. Short
bases: bootparam_40's two terminal primes, bootparam_30's two main primes,
rns24's one main prime; fused and separate key inner products; BaseConv
fused into the transform, on the FP64 and on the int8 matrix cores. Also the
level -1 ModUp / ModDown entry points and ck_context_set_short_base's
argument checks. (ModUpToMax as a whole, with these switches on both sides,
is compared word for word inside tests/cpp/test_boot.cpp.)"""
import numpy as np
import pytest

from conftest import SEED_BASE, load_params
from gpu_util import assert_same, uniform

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ck():
    import cheddar_amd
    return cheddar_amd


@pytest.fixture(scope="module", params=["bootparam_40.json", "bootparam_30.json",
                                        "rns24_aux8.json"])
def base(request, ck):
    import ckks_oracle as O
    d = load_params(request.param)
    P = O.Param.from_json(d)
    ctx = ck.Context.from_json(d)
    nm, nt = P.level_np(0)
    ctx.set_short_base(nm, nt)
    return dict(name=request.param, P=P, ctx=ctx, nm=nm, nt=nt)


def short_key(P, nm, nt, seed):
    """a uniform key block over the short base's limbs, then as many
    auxiliary limbs"""
    k = nm + nt
    pidx = P.np_primes(nm, nt) + [P.L + j for j in range(k)]
    return uniform(P, pidx, seed), uniform(P, pidx, seed + 1)


def test_short_base_shape(ck, base):
    ctx, nm, nt = base["ctx"], base["nm"], base["nt"]
    assert ctx.level_np(-1) == (nm, nt)
    assert ctx.modup_beta(-1) == 1
    ctx.set_short_base(nm, nt)  # the same base again: a no-op


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("bconv", [False, True, "int8"])
def test_mult_key_level_minus_one(ck, base, fused, bconv):
    P, ctx, nm, nt = base["P"], base["ctx"], base["nm"], base["nt"]
    pidx = P.np_primes(nm, nt)
    print(f"short-base MultKey {base['name']} fused={fused} bconv={bconv}", flush=True)
    seed = SEED_BASE + 9 * 1000 + 10 * len(base["name"]) + 2 * int(fused)
    ct = [uniform(P, pidx, seed + j) for j in range(2)]
    for j in range(2):  # extremes on both polynomials
        for i, p in enumerate(pidx):
            q = P.prime(p)
            ct[j][i, :4] = (0, q - 1, (q - 1) // 2, (q + 1) // 2)
    kb, ka = short_key(P, nm, nt, seed + 100)
    want = P.mult_key_short_base(nm, nt, ct, (kb, ka))
    ctx.set_fused_key_rows(fused)
    ctx.set_bconv_mfma(bconv)
    try:
        d_ct = [ck.to_device(v) for v in ct]
        out = [torch.full_like(d_ct[0], -1) for _ in range(2)]
        ctx.mult_key(out, d_ct, ([ck.to_device(kb)], [ck.to_device(ka)]), -1)
        torch.cuda.synchronize()
        for j in range(2):
            assert_same(ck.to_host(out[j]), want[j], f"short-base MultKey poly {j}")
        # in place on the input ciphertext
        ctx.mult_key(d_ct, d_ct, ([ck.to_device(kb)], [ck.to_device(ka)]), -1)
        torch.cuda.synchronize()
        for j in range(2):
            assert_same(ck.to_host(d_ct[j]), want[j], f"in-place short-base MultKey poly {j}")
    finally:
        ctx.set_fused_key_rows(True)
        ctx.set_bconv_mfma(None)


def test_modup_moddown_level_minus_one(ck, base):
    """ModUp at level -1 raises to the short base's auxiliary primes (its own
    limbs copied), ModDown at level -1 removes them: ModDown(P' x) = x"""
    P, ctx, nm, nt = base["P"], base["ctx"], base["nm"], base["nt"]
    k = nm + nt
    pidx = P.np_primes(nm, nt)
    x = uniform(P, pidx, SEED_BASE + 9 * 1000 + 500)
    d_x = ck.to_device(x)
    up = torch.full((2 * k, P.n), -1, dtype=torch.int32, device="cuda")
    ctx.modup(up, d_x, -1)
    torch.cuda.synchronize()
    got = ck.to_host(up)
    assert_same(got[:k], x, "ModUp(-1) own limbs")
    # P' x on the q limbs, 0 on the auxiliary limbs: ModDown gives x exactly
    pp = []
    for p in pidx:
        q = P.prime(p)
        v = 1
        for j in range(k):
            v = v * P.prime(P.L + j) % q
        pp.append(v * (1 << 32) % q)
    px = np.zeros((2 * k, P.n), dtype=np.uint32)
    px[:k] = P.mult_const(x, np.array(pp, dtype=np.uint32), pidx)
    d_px = ck.to_device(px)
    down = torch.full((k, P.n), -1, dtype=torch.int32, device="cuda")
    ctx.moddown([down], [d_px], -1, kind=0)
    torch.cuda.synchronize()
    assert_same(ck.to_host(down), x, "ModDown(-1)(P' x)")


def test_short_base_argument_checks(ck, base):
    P, ctx, nm, nt = base["P"], base["ctx"], base["nm"], base["nt"]
    with pytest.raises(ck.CheddarError):
        ctx.set_short_base(nm + 1, nt)  # a different base once one is set
    k = nm + nt
    x = ck.to_device(uniform(P, P.np_primes(nm, nt), SEED_BASE + 9 * 1000 + 600))
    up = torch.empty((2 * k, P.n), dtype=torch.int32, device="cuda")
    with pytest.raises(ck.CheddarError):
        ctx.modup(up, x, -2)
    with pytest.raises(ck.CheddarError):
        ctx.moddown([x], [up], -1, kind=1)  # no rescale at level -1


def test_short_base_limits(ck, base):
    """1..alpha primes within level 0; level -1 is refused until set up"""
    P, nm, nt = base["P"], base["nm"], base["nt"]
    ctx = ck.Context.from_json(load_params(base["name"]))
    with pytest.raises(ck.CheddarError):
        ctx.set_short_base(P.alpha + 1, 0)
    with pytest.raises(ck.CheddarError):
        ctx.set_short_base(0, 0)
    assert ctx.modup_beta(-1) == -1
    x = ck.to_device(uniform(P, P.np_primes(nm, nt), SEED_BASE + 9 * 1000 + 700))
    with pytest.raises(ck.CheddarError):
        ctx.mult_key([x, x], [x, x], ([x], [x]), -1)
