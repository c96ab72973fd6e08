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
