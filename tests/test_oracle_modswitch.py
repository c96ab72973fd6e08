"""Oracle pins for BaseConv, ModUp, ModDown, Rescale and ModDownAndRescale
(CPU only), by big-integer CRT reconstruction.

Reference semantics (src/core/ModSwitch.cu:23-684): fast base conversion
y = sum_i [x * s^_i^-1]_{s_i} (centred) * s^_i, ModUp lifts each dnum group
of limbs to the whole basis, ModDown / Rescale compute (a - E) / D where E is
the fast base conversion of a's residues modulo the dropped primes D (with
the BitPacker restore factor for terminal-prime level changes). The checks
reconstruct coefficients with the Chinese remainder theorem and verify:
  * the lifted integers agree with the input modulo the group primes and are
    bounded by |group| * Q_g / 2 (ModUp);
  * out * D = a * restore - E exactly, with E the textbook fast base
    conversion, and |E| <= len(D) * D / 2 (ModDown family);
  * the exact identity ModDown(P * b || 0) = b.
"""
import json
import os
import random

import numpy as np
import pytest

from conftest import GOLDEN, SEED_BASE

import ckks_oracle as O

with open(os.path.join(GOLDEN, "vectors.json")) as f:
    GOLD = json.load(f)


def centred(x, m):
    x %= m
    return x - m if x > m // 2 else x


def crt(res, primes):
    M = 1
    for p in primes:
        M *= p
    x = 0
    for r, p in zip(res, primes):
        Mi = M // p
        x += int(r) * Mi * pow(Mi, -1, p)
    return centred(x, M), M


def coeffs(P, data, nm, nt, na):
    """NTT-domain Montgomery-form limbs -> plain coefficient residues"""
    return P.intt_np(data, nm, nt, na, mont_conv=True)


def fast_bconv(x, src):
    """textbook fast base conversion lift of integer x from basis src"""
    S = 1
    for s in src:
        S *= s
    return sum(centred(x * pow(S // s, -1, s), s) * (S // s) for s in src)


def test_bconv_known_answers(oracle_rns24):
    P = oracle_rns24
    for c in GOLD["bconv"]:
        idx = {int(P.prime(i)): i for i in range(P.num_primes)}
        sp = [idx[s] for s in c["src"]]
        dp = [idx[t] for t in c["dst"]]
        _, table = P.bconv_tables(sp, dp)
        src = np.tile(np.array(c["v"], dtype=np.int32)[:, None], (1, P.n))
        y = P.bconv(src, sp, dp, table)
        assert [int(y[m, 0]) for m in range(len(dp))] == c["y"]
        assert np.all(y == y[:, :1])


def test_bconv_tables_definition(oracle_rns24):
    P = oracle_rns24
    sp, dp = list(range(24, 32)), list(range(0, 24))
    c1, table = P.bconv_tables(sp, dp, restore_pidx=[3, 7])
    S = 1
    for i in sp:
        S *= P.prime(i)
    for i, si in enumerate(sp):
        s = P.prime(si)
        want = pow((S // s) * P.n, -1, s) * P.prime(3) * P.prime(7) % s
        assert int(c1[i]) == want
        for m, ti in enumerate(dp):
            t = P.prime(ti)
            assert int(table[m, i]) == centred((S // s) * 2 ** 64, t)


@pytest.mark.parametrize("level", [23, 10, 0])
def test_modup_crt(oracle_rns24, level):
    P = oracle_rns24
    nq = P.num_q(level)
    qp = P.level_primes(level)
    allp = P.level_primes(level, with_aux=True)
    x = P.uniform(qp, SEED_BASE + 50 + level)
    xc = coeffs(P, x, nq, 0, 0)
    out = P.modup(x, level)
    rng = random.Random(level)
    ks = [0, P.n - 1] + [rng.randrange(P.n) for _ in range(4)]
    for g in range(out.shape[0]):
        s0, s1 = g * P.alpha, min((g + 1) * P.alpha, nq)
        # own limbs are copied unchanged
        assert np.array_equal(out[g, s0:s1], x[s0:s1])
        yc = coeffs(P, out[g], nq, 0, P.alpha)
        Qg = 1
        for t in range(s0, s1):
            Qg *= P.prime(qp[t])
        for k in ks:
            X, _ = crt([yc[t, k] for t in range(len(allp))], [P.prime(i) for i in allp])
            for t in range(s0, s1):
                assert (X - int(xc[t, k])) % P.prime(qp[t]) == 0
            assert abs(X) <= (s1 - s0) * Qg // 2
            assert X == fast_bconv(crt([xc[t, k] for t in range(s0, s1)],
                                       [P.prime(qp[t]) for t in range(s0, s1)])[0],
                                   [P.prime(qp[t]) for t in range(s0, s1)])


def moddown_check(P, level, kind, a_ntt, out_ntt, restore=1):
    """out * D == a * restore - E (E = fast BaseConv lift of a over D)"""
    nm, nt = P.level_np(level)
    src_aux = 0 if kind == O.OR_RESCALE else P.alpha
    sp = P.np_primes(nm, nt, src_aux)
    if kind == O.OR_MODDOWN:
        dm, dt = nm, nt
    else:
        dm, dt = P.level_np(level - 1)
    dp = P.np_primes(dm, dt, 0)
    ac = coeffs(P, a_ntt, nm, nt, src_aux)
    oc = coeffs(P, out_ntt, dm, dt, 0)
    removed = [p for p in sp if p not in dp]
    D = 1
    for p in removed:
        D *= P.prime(p)
    rng = random.Random(level * 3 + kind)
    for k in [0, P.n - 1] + [rng.randrange(P.n) for _ in range(5)]:
        A, M = crt([ac[t, k] for t in range(len(sp))], [P.prime(p) for p in sp])
        C, Mq = crt([oc[t, k] for t in range(len(dp))], [P.prime(p) for p in dp])
        E = fast_bconv(A * restore, [P.prime(p) for p in removed])
        assert abs(E) <= len(removed) * D // 2
        for t, p in enumerate(dp):
            q = P.prime(p)
            assert (C * D - (A * restore - E)) % q == 0, (level, kind, k, t)


@pytest.mark.parametrize("level,kind", [(23, 0), (23, 1), (23, 2), (7, 2), (1, 1), (0, 0)])
def test_moddown_family_crt(oracle_rns24, level, kind):
    P = oracle_rns24
    src = P.uniform(P.level_primes(level, with_aux=(kind != O.OR_RESCALE)),
                    SEED_BASE + 60 + level + 7 * kind)
    out = P.moddown(src, level, kind)
    moddown_check(P, level, kind, src, out)


def test_moddown_exact_identity(oracle_rns24):
    """ModDown(P * b || 0) == b exactly"""
    P = oracle_rns24
    level = 23
    qp = P.level_primes(level)
    b = P.uniform(qp, SEED_BASE + 70)
    pp = np.array([P.p_prod_mont(i) for i in qp], dtype=np.uint32)
    pb = P.mult_const(b, pp, qp)
    a = np.concatenate([pb, np.zeros((P.alpha, P.n), np.uint32)])
    assert np.array_equal(P.moddown(a, level, O.OR_MODDOWN), b)


def test_rescale_is_rounding(oracle_rns24):
    """Rescale = round(a / q_last) on every coefficient"""
    P = oracle_rns24
    level = 5
    qp = P.level_primes(level)
    a = P.uniform(qp, SEED_BASE + 71)
    out = P.moddown(a, level, O.OR_RESCALE)
    ac = coeffs(P, a, 6, 0, 0)
    oc = coeffs(P, out, 5, 0, 0)
    ql = P.prime(qp[-1])
    for k in (0, 1, 777, P.n - 1):
        A, _ = crt([ac[t, k] for t in range(6)], [P.prime(p) for p in qp])
        C, _ = crt([oc[t, k] for t in range(5)], [P.prime(p) for p in qp[:5]])
        assert C == (A - centred(A, ql)) // ql


# bootparam_40: terminal primes with BitPacker restore on level changes
@pytest.mark.parametrize("level,kind", [(28, 0), (28, 1), (28, 2), (3, 1), (3, 2), (2, 1), (2, 2), (1, 2)])
def test_boot40_moddown_crt(oracle_boot40, level, kind):
    P = oracle_boot40
    src = P.uniform(P.level_primes(level, with_aux=(kind != O.OR_RESCALE)),
                    SEED_BASE + 80 + level + 7 * kind)
    out = P.moddown(src, level, kind)
    restore = 1
    if kind != O.OR_MODDOWN:
        nm, nt = P.level_np(level)
        dm, dt = P.level_np(level - 1)
        new = [p for p in P.np_primes(dm, dt, 0) if p not in P.np_primes(nm, nt, 0)]
        for p in new:
            restore *= P.prime(p)
    moddown_check(P, level, kind, src, out, restore)


def test_boot40_modup_crt(oracle_boot40):
    P = oracle_boot40
    level = 2  # (1 main, 4 terminal): terminal padding shifts the groups
    nm, nt = P.level_np(level)
    nq = nm + nt
    qp = P.level_primes(level)
    allp = P.level_primes(level, with_aux=True)
    x = P.uniform(qp, SEED_BASE + 90)
    xc = coeffs(P, x, nm, nt, 0)
    out = P.modup(x, level)
    yc = coeffs(P, out[0], nm, nt, P.alpha)
    for k in (0, 5, P.n - 1):
        X, _ = crt([yc[t, k] for t in range(len(allp))], [P.prime(i) for i in allp])
        src = [P.prime(qp[t]) for t in range(nq)]
        assert X == fast_bconv(crt([xc[t, k] for t in range(nq)], src)[0], src)
