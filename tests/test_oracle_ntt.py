"""Oracle pins for the negacyclic NTT / INTT (CPU only).

The transform definition (src/core/NTT.cu:231-454, NTTUtils.cuh:19-58) is
pinned by direct evaluation out[k] = sum_j a_j psi^((2 brv(k) + 1) j) (big
integer known answers at N = 16 and sampled positions at N = 2^16) and by the
negacyclic convolution theorem against schoolbook products -- the CPU
workload of the reference configuration: N = 2^16, 24 limbs, 64 random pairs.
"""
import json
import os
import random

import numpy as np

from conftest import GOLDEN, SEED_BASE

import ckks_oracle as O

with open(os.path.join(GOLDEN, "vectors.json")) as f:
    GOLD = json.load(f)


def brv(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2)


def test_ntt_known_answers_n16():
    for c in GOLD["ntt16"]:
        q = c["q"]
        P = O.Param(c["log_n"], [q], [2146959361] if q != 2146959361 else [2146041857],
                    level_config=[[1, 0]])
        a = np.array([c["in"]], dtype=np.uint32)
        assert P.ntt(a, [0]).tolist()[0] == c["out"]


def test_ntt_direct_evaluation_n65536(oracle_rns24):
    P = oracle_rns24
    rng = random.Random(11)
    for pi in (0, 23, 31):
        q, psi = P.prime(pi), P.psi(pi)
        a = P.uniform([pi], SEED_BASE + 20 + pi)
        A = P.ntt(a, [pi])
        aj = [int(v) for v in a[0]]
        for k in [0, 1, 65535] + [rng.randrange(65536) for _ in range(2)]:
            w = pow(psi, 2 * brv(k, 16) + 1, q)
            acc, p = 0, 1
            for j in range(P.n):
                acc += aj[j] * p
                p = p * w % q
            assert int(A[0, k]) == acc % q, (pi, k)


def test_roundtrip_24_limbs(oracle_rns24):
    P = oracle_rns24
    x = P.uniform(P.level_primes(23), SEED_BASE + 30)
    X = P.ntt_np(x, 24, 0)
    assert np.array_equal(P.intt_np(X, 24, 0, mont_conv=False), x)
    # the Montgomery-leaving INTT multiplies by 2^-32 as well
    pidx = P.level_primes(23)
    y = P.intt_np(X, 24, 0, mont_conv=True)
    for t in (0, 11, 23):
        q = P.prime(pidx[t])
        rinv = pow(2 ** 32, -1, q)
        assert np.array_equal(y[t], (x[t].astype(object) * rinv % q).astype(np.uint32))


def test_mont_conversion_and_linearity(oracle_rns24):
    P = oracle_rns24
    pidx = P.level_primes(23)[:4]
    x = P.uniform(pidx, SEED_BASE + 31)
    y = P.uniform(pidx, SEED_BASE + 32)
    q = np.array([P.prime(i) for i in pidx], dtype=np.uint64)[:, None]
    xr = (x.astype(np.uint64) * ((1 << 32) % q) % q).astype(np.uint32)
    assert np.array_equal(P.ntt(x, pidx, mont_conv=True), P.ntt(xr, pidx))
    s = ((x.astype(np.uint64) * 3 + y.astype(np.uint64) * 5) % q).astype(np.uint32)
    lhs = P.ntt(s, pidx).astype(np.uint64)
    rhs = (P.ntt(x, pidx).astype(np.uint64) * 3 + P.ntt(y, pidx).astype(np.uint64) * 5) % q
    assert np.array_equal(lhs, rhs)


def negacyclic_coeff(a, b, k, q):
    """coefficient k of a*b mod (X^N + 1, q), products reduced in int64"""
    a = a.astype(np.int64)
    b = b.astype(np.int64)
    pos = int(((a[: k + 1] * b[k::-1]) % q).sum() % q)
    neg = int(((a[k + 1:] * b[:k:-1]) % q).sum() % q) if k + 1 < len(a) else 0
    return (pos - neg) % q


def negacyclic_full(a, b, q):
    """full product via Kronecker substitution on Python big integers"""
    n = len(a)
    slot = 80
    A = int.from_bytes(b"".join(int(v).to_bytes(10, "little") for v in a), "little")
    B = int.from_bytes(b"".join(int(v).to_bytes(10, "little") for v in b), "little")
    C = (A * B).to_bytes(20 * n + 10, "little")
    mask = slot // 8
    coef = [int.from_bytes(C[i * mask:(i + 1) * mask], "little") for i in range(2 * n)]
    return np.array([(coef[i] - coef[i + n]) % q for i in range(n)], dtype=np.uint64)


def test_convolution_theorem_64_pairs(oracle_rns24):
    """NTT-domain pointwise product == negacyclic schoolbook product, for 64
    random pairs spread over the 24 limbs (sampled output coefficients for
    all pairs, full products for two of them)."""
    P = oracle_rns24
    pidx = P.level_primes(23)
    rng = random.Random(12)
    for pair in range(64):
        t = pair % 24
        pi = pidx[t]
        q = P.prime(pi)
        a = P.uniform([pi], SEED_BASE + 1000 + 2 * pair)
        b = P.uniform([pi], SEED_BASE + 1001 + 2 * pair)
        A, B = P.ntt(a, [pi]), P.ntt(b, [pi])
        C = (A.astype(np.uint64) * B.astype(np.uint64) % q).astype(np.uint32)
        c = P.intt_raw(C, [pi]).astype(np.uint64) * pow(P.n, -1, q) % q
        if pair < 2:
            assert np.array_equal(c[0], negacyclic_full(a[0], b[0], q))
        for k in [0, P.n - 1] + [rng.randrange(P.n) for _ in range(4)]:
            assert int(c[0, k]) == negacyclic_coeff(a[0], b[0], k, q), (pair, k)


def test_intt_mult_const_semantics(oracle_rns24):
    """INTTAndMultConst: out = N * INTT(x) * c * 2^-32, canonical or centred"""
    P = oracle_rns24
    pidx = P.level_primes(23)[:3]
    x = P.uniform(pidx, SEED_BASE + 33)
    consts = [12345, 2 ** 30 + 7, 1]
    true_inv = P.intt_np(x, 3, 0, mont_conv=False)  # exact inverse transform
    for normalize in (False, True):
        out = P.intt_mult_const(x, pidx, consts, normalize)
        for t, pi in enumerate(pidx):
            q = P.prime(pi)
            f = P.n * consts[t] * pow(2 ** 32, -1, q) % q
            want = true_inv[t].astype(object) * f % q
            got = out[t].astype(np.int64)
            if normalize:
                assert got.max() <= (q - 1) // 2 and got.min() >= -(q - 1) // 2
            assert np.array_equal(got % q, want.astype(np.int64))
