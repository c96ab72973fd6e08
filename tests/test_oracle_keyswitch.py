"""Oracle pins for Tensor, Permute and the key-switching compositions (CPU).

Semantic round trips with real keys (the reference's own test semantics,
unittest/BasicTest.cpp:327-419, at full N = 2^16): encrypt -> HMult
(relinearize + rescale) -> decrypt equals round(m1 m2 / q_last) up to the
ModDown error, HRot by r decrypts to m(X^(5^r)), HConj to m(X^-1); plus
closed-form checks of Tensor and of the automorphism index map.
"""
import random

import numpy as np
import pytest

from conftest import SEED_BASE
from ckks_keys import Keys, automorph_coef, negacyclic_product

import ckks_oracle as O


def test_tensor_is_product(oracle_rns24):
    """Karatsuba Tensor == schoolbook (b1 + a1 S)(b2 + a2 S) coefficients"""
    P = oracle_rns24
    pidx = P.level_primes(23)[:3]
    ins = [P.uniform(pidx, SEED_BASE + 300 + i) for i in range(4)]
    d0, d1, d2 = P.tensor(*ins, pidx)
    b1, a1, b2, a2 = ins
    assert np.array_equal(d0, P.mult(b1, b2, pidx))
    assert np.array_equal(d2, P.mult(a1, a2, pidx))
    assert np.array_equal(d1, P.add(P.mult(b1, a2, pidx), P.mult(a1, b2, pidx), pidx))


def test_permute_is_automorphism(oracle_rns24):
    """Permute(NTT(m)) == NTT(m(X^g)) for rotations and conjugation, and
    sigma_g o sigma_h = sigma_gh"""
    P = oracle_rns24
    pidx = P.level_primes(23)[:2]
    rng = np.random.default_rng(5)
    m = rng.integers(-1000, 1000, size=P.n)
    q = np.array([P.prime(i) for i in pidx], dtype=np.int64)[:, None]
    M = P.ntt(((m[None, :] % q)).astype(np.uint32), pidx)
    for rot in (1, 3, 1234, -1):
        g = P.galois(rot)
        want_c = np.array(automorph_coef(m, g), dtype=np.int64)
        want = P.ntt((want_c[None, :] % q).astype(np.uint32), pidx)
        assert np.array_equal(P.permute(M, g), want), rot
    g1, g2 = P.galois(7), P.galois(11)
    assert np.array_equal(P.permute(P.permute(M, g1), g2),
                          P.permute(M, g1 * g2 % (2 * P.n)))
    conj = P.galois(-1)
    assert np.array_equal(P.permute(P.permute(M, conj), conj), M)


@pytest.fixture(scope="module")
def keys24(oracle_rns24):
    return Keys(oracle_rns24, SEED_BASE + 400)


@pytest.mark.parametrize("level", [23, 6])
def test_hmult_decrypts_to_product(oracle_rns24, keys24, level):
    P = oracle_rns24
    K = keys24
    rng = np.random.default_rng(level)
    m1 = rng.integers(-2 ** 24, 2 ** 24, size=P.n)
    m2 = rng.integers(-2 ** 24, 2 ** 24, size=P.n)
    ct1 = K.encrypt(m1, level)
    ct2 = K.encrypt(m2, level)
    evk = K.mult_key()
    out = P.hmult(level, ct1, ct2, evk, rescale=True)
    ql = P.prime(P.level_primes(level)[-1])
    prod = negacyclic_product(m1, m2)
    ks = [0, 1, P.n - 1] + [int(v) for v in rng.integers(0, P.n, size=5)]
    got = K.decrypt(out, level - 1, ks)
    for k, v in zip(ks, got):
        want = round(prod[k] / ql)
        assert abs(v - want) < 2 ** 14, (k, v, want)
    # relinearize without rescale keeps the product at scale q_last
    out2 = P.hmult(level, ct1, ct2, evk, rescale=False)
    got2 = K.decrypt(out2, level, ks[:3])
    for k, v in zip(ks[:3], got2):
        assert abs(v - prod[k]) < 2 ** 45


@pytest.mark.parametrize("rot", [1, 1234, 32767])
def test_hrot_decrypts_to_rotation(oracle_rns24, keys24, rot):
    P = oracle_rns24
    K = keys24
    level = 12
    rng = np.random.default_rng(rot)
    m = rng.integers(-2 ** 24, 2 ** 24, size=P.n)
    ct = K.encrypt(m, level)
    out = P.hrot(level, ct, K.rot_key(rot), rot)
    want = automorph_coef(m, P.galois(rot))
    ks = [0, 1, P.n - 1] + [int(v) for v in rng.integers(0, P.n, size=5)]
    for k, v in zip(ks, K.decrypt(out, level, ks)):
        assert abs(v - want[k]) < 2 ** 14, (rot, k)


def test_hconj_decrypts_to_conjugate(oracle_rns24, keys24):
    """HConj = MultKey + Permute(2N - 1) (Context.cpp:793-798)"""
    P = oracle_rns24
    K = keys24
    level = 9
    rng = np.random.default_rng(99)
    m = rng.integers(-2 ** 24, 2 ** 24, size=P.n)
    b, a = K.encrypt(m, level)
    kb, ka = K.conj_key()
    acc_b, acc_a = P.mult_key_no_moddown(level, a, kb, ka, add_b=b)
    rb = P.moddown(acc_b, level, O.OR_MODDOWN)
    ra = P.moddown(acc_a, level, O.OR_MODDOWN)
    g = P.galois(-1)
    out = (P.permute(rb, g), P.permute(ra, g))
    got = P.hconj(level, (b, a), (kb, ka))
    assert np.array_equal(got[0], out[0]) and np.array_equal(got[1], out[1])
    want = automorph_coef(m, g)
    ks = [0, 1, 2, P.n - 1, 4321]
    for k, v in zip(ks, K.decrypt(out, level, ks)):
        assert abs(v - want[k]) < 2 ** 14, k


def test_boot40_hmult_decrypts(oracle_boot40):
    """terminal primes + restore: HMult at the top level of bootparam_40
    (43 main + 4 terminal limbs, dnum 4) decrypts to the rescaled product"""
    P = oracle_boot40
    K = Keys(P, SEED_BASE + 500)
    level = P.max_level
    rng = np.random.default_rng(3)
    m1 = rng.integers(-2 ** 24, 2 ** 24, size=P.n)
    m2 = rng.integers(-2 ** 24, 2 ** 24, size=P.n)
    out = P.hmult(level, K.encrypt(m1, level), K.encrypt(m2, level), K.mult_key())
    nm, nt = P.level_np(level)
    dm, dt = P.level_np(level - 1)
    removed = [p for p in P.np_primes(nm, nt) if p not in P.np_primes(dm, dt)]
    D = 1
    for p in removed:
        D *= P.prime(p)
    prod = negacyclic_product(m1, m2)
    ks = [0, 17, P.n - 1]
    for k, v in zip(ks, K.decrypt(out, level - 1, ks)):
        assert abs(v - round(prod[k] / D)) < 2 ** 16, k


@pytest.mark.parametrize("rot", [1, 1234, 32767])
def test_hoisted_baby_step_is_permuted_key_product(oracle_rns24, rot):
    """BSFusedKernel's scatter by 5^(N/2 - r) (Hoist.cu:21-78) is the
    Permute gather by 5^r of MultKeyNoModDown(a) + P b, on all Q*P limbs"""
    from gpu_util import full_key
    P = oracle_rns24
    level = 9
    pidx = P.level_primes(level)
    ct = [P.uniform(pidx, SEED_BASE + 700 + i) for i in range(2)]
    key = full_key(P, SEED_BASE + 710)
    bs = P.hoisted_baby_step(level, ct, key, rot)
    acc = P.mult_key_no_moddown(level, ct[1], key[0], key[1], add_b=ct[0])
    g = P.galois(rot)
    assert np.array_equal(bs[0], P.permute(acc[0], g))
    assert np.array_equal(bs[1], P.permute(acc[1], g))


@pytest.mark.parametrize("level,rot", [(23, 5), (4, 20000)])
def test_hoisted_then_moddown_equals_hrot(oracle_rns24, level, rot):
    """ModDown commutes with the automorphism and ModDown(P b + s) =
    b + ModDown(s): the hoisted baby step followed by ModDown is HRot"""
    from gpu_util import full_key
    P = oracle_rns24
    pidx = P.level_primes(level)
    ct = [P.uniform(pidx, SEED_BASE + 720 + i) for i in range(2)]
    key = full_key(P, SEED_BASE + 730)
    bs = P.hoisted_baby_step(level, ct, key, rot)
    got = [P.moddown(v, level, O.OR_MODDOWN) for v in bs]
    want = P.hrot(level, ct, key, rot)
    assert np.array_equal(got[0], want[0])
    assert np.array_equal(got[1], want[1])


# ---- level -1: the dense -> sparse switch of sparse-secret encapsulation ----

def test_short_base_mult_key_equals_level_switch_where_they_coincide(oracle_rns24):
    """rns24's level 7 has 8 main primes = alpha and no terminal primes, so
    its ModSwitchHandler is a single unpadded group raised to all 8
    auxiliary primes: exactly the level -1 handler of a short base of those
    8 primes (ModSwitch.cu:175-232). MultKey there (Context.cpp:584-607)
    and the short-base MultKey agree word for word."""
    from gpu_util import full_key
    P = oracle_rns24
    level = 7
    nm, nt = P.level_np(level)
    assert (nm, nt) == (8, 0) and P.alpha == 8 and P.modup_beta(level) == 1
    pidx = P.level_primes(level)
    ct = [P.uniform(pidx, SEED_BASE + 740 + i) for i in range(2)]
    kb, ka = full_key(P, SEED_BASE + 750)
    acc_b, acc_a = P.mult_key_no_moddown(level, ct[1], kb, ka, add_b=ct[0])
    want = [P.moddown(v, level, O.OR_MODDOWN) for v in (acc_b, acc_a)]
    rows = pidx + [P.L + j for j in range(8)]
    got = P.mult_key_short_base(nm, nt, ct, (kb[0][rows], ka[0][rows]))
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def dts_key(P, dense, sparse, nm, nt, err=3):
    """the dense -> sparse key over the short base and its auxiliary primes
    (UserInterface.cu PrepareEvk for kDenseToSparseKeyIndex): b = -a s' + e
    + P' s on the short base's limbs, -a s' + e on the auxiliary limbs"""
    qp = P.np_primes(nm, nt)
    k = len(qp)
    pidx = qp + [P.L + j for j in range(k)]
    a = dense.uniform(pidx)
    e = dense.ntt_small(dense.small(err), pidx)
    b = P.sub(e, P.mult(a, sparse.s[pidx], pidx), pidx)
    pp = []
    for p in qp:
        q = P.prime(p)
        v = 1
        for j in range(k):
            v = v * P.prime(P.L + j) % q
        pp.append(v * (1 << 32) % q)
    b[:k] = P.add(b[:k], P.mult_const(dense.s[qp], np.array(pp, dtype=np.uint32), qp), qp)
    return b, a


@pytest.mark.parametrize("name", ["bootparam_40.json", "bootparam_30.json"])
def test_dts_mult_key_switches_dense_to_sparse(name):
    """BootContext::ModUpToMax's dense -> sparse switch (BootContext.cpp:
    205-210 -> Context::MultKey at level -1): a level-0 encryption under the
    dense secret decrypts under the sparse secret (Hamming weight 32) after
    the switch, up to the key-switching error. bootparam_40's short base is
    two terminal primes, bootparam_30's two main primes."""
    from conftest import load_params
    P = O.Param.from_json(load_params(name))
    nm, nt = P.level_np(0)
    dense = Keys(P, SEED_BASE + 760)
    sparse = Keys(P, SEED_BASE + 761, hamming=32)
    key = dts_key(P, dense, sparse, nm, nt)
    rng = np.random.default_rng(7)
    m = rng.integers(-2 ** 30, 2 ** 30, size=P.n)  # |m| < Q / 2^9
    ct = dense.encrypt(m, 0)
    out = P.mult_key_short_base(nm, nt, ct, key)
    ks = [0, 1, P.n - 1] + [int(v) for v in rng.integers(0, P.n, size=5)]
    for k, v in zip(ks, sparse.decrypt(out, 0, ks)):
        assert abs(v - m[k]) < 2 ** 16, (k, v, m[k])
    # the switched ciphertext does not decrypt under the dense secret
    # (bootparam_30's short base is two ~2^20 primes: Q ~ 2^40)
    Q = 1
    for p in P.np_primes(nm, nt):
        Q *= P.prime(p)
    bad = dense.decrypt(out, 0, ks[:3])
    assert any(abs(v - m[k]) > Q >> 12 for k, v in zip(ks[:3], bad))


def std_key(P, dense, sparse, err=3):
    """the sparse -> dense key at the top level, one block over the top
    level's q limbs and the alpha auxiliary limbs: b = -a s + e + P s' on
    the q limbs (the whole of Q as one digit: the raised a is a lift of a
    short-base value, so a e / P stays small)"""
    tq = P.level_primes(P.max_level)
    pidx = tq + [P.L + j for j in range(P.alpha)]
    a = dense.uniform(pidx)
    e = dense.ntt_small(dense.small(err), pidx)
    b = P.sub(e, P.mult(a, dense.s[pidx], pidx), pidx)
    pp = np.array([P.p_prod_mont(p) for p in tq], dtype=np.uint32)
    nq = len(tq)
    b[:nq] = P.add(b[:nq], P.mult_const(sparse.s[tq], pp, tq), tq)
    return b, a


@pytest.mark.parametrize("name,sse", [("bootparam_40.json", False),
                                      ("bootparam_40.json", True),
                                      ("bootparam_30.json", True)])
def test_mod_up_to_max_lifts_by_multiples_of_q0(name, sse):
    """BootContext::ModUpToMax (BootContext.cpp:186-268): the raised
    ciphertext decrypts at the top level to m + e + Q0 I with |I| at most
    the sparse Hamming weight + 2 (the lift of b + a s' from the short base
    Q0), under the sparse secret without encapsulation and under the dense
    secret after the dense -> sparse -> dense switches (then e also holds
    the two key-switching errors)."""
    from conftest import load_params
    P = O.Param.from_json(load_params(name))
    nm, nt = P.level_np(0)
    dense = Keys(P, SEED_BASE + 770)
    sparse = Keys(P, SEED_BASE + 771, hamming=32)
    owner = dense if sse else sparse
    rng = np.random.default_rng(11)
    m = rng.integers(-2 ** 20, 2 ** 20, size=P.n)
    ct = owner.encrypt(m, 0)
    if sse:
        out = P.mod_up_to_max_ct(nm, nt, ct, dts_key(P, dense, sparse, nm, nt),
                                 std_key(P, dense, sparse))
    else:
        out = P.mod_up_to_max_ct(nm, nt, ct)
    Q0 = 1
    for p in P.np_primes(nm, nt):
        Q0 *= P.prime(p)
    ks = [0, 1, P.n - 1] + [int(v) for v in rng.integers(0, P.n, size=13)]
    lifts = []
    for k, v in zip(ks, owner.decrypt(out, P.max_level, ks)):
        d = v - int(m[k])
        r = (d + Q0 // 2) % Q0 - Q0 // 2
        I = (d - r) // Q0
        assert abs(r) < (2 ** 20 if sse else 2 ** 6), (k, r)
        assert abs(I) <= 34, (k, I)
        lifts.append(I)
    assert any(I != 0 for I in lifts)
