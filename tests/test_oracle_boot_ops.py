"""Oracle pins for the element-wise operations that bootstrapping adds
(AddConst / SubConst / SubOppositeConst, MultImaginaryUnit, ModUpToMax;
src/core/ElementWise.cu:145-223, 579-658), each against what the
mathematics fixes rather than a restatement of the formula:
- X^(N/2) applied twice is -1, and equals the negacyclic shift by N/2 in
  the coefficient domain carried through the forward NTT;
- ModUpToMax of the CRT residues [x (q0 q1 / q_i)^-1]_{q_i} of an integer
  x in (-q0 q1 / 2, q0 q1 / 2] is x mod t for every destination prime t
  (Python integers);
- add / sub / opposite-sub of constants are inverse to each other and to
  negation.
"""
import numpy as np

from conftest import SEED_BASE


def test_imaginary_unit_squares_to_minus_one(oracle_rns24):
    P = oracle_rns24
    pidx = P.np_primes(24, 0, 8)
    x = P.uniform(pidx, SEED_BASE + 4000)
    ix = P.mult_imaginary_unit(x, pidx)
    assert np.array_equal(P.mult_imaginary_unit(ix, pidx), P.neg(x, pidx))
    assert not np.array_equal(ix, x)


def test_imaginary_unit_is_negacyclic_shift(oracle_rns24):
    P = oracle_rns24
    n, half = P.n, P.n // 2
    pidx = [0, 5, 31]
    c = P.uniform(pidx, SEED_BASE + 4001)
    shifted = np.empty_like(c)
    for t, pi in enumerate(pidx):
        q = P.prime(pi)
        # c(X) * X^(N/2): coefficient j moves to j + N/2, wrapping ones negate
        shifted[t, half:] = c[t, :half]
        shifted[t, :half] = (q - c[t, half:].astype(np.int64)) % q
    want = P.ntt(shifted, pidx)
    got = P.mult_imaginary_unit(P.ntt(c, pidx), pidx)
    assert np.array_equal(got, want)


def test_modup_to_max_is_crt_lift(oracle_boot40):
    P = oracle_boot40
    base = P.np_primes(0, 2)  # the short base of bootparam_40: two terminal primes
    q0, q1 = P.prime(base[0]), P.prime(base[1])
    qq = q0 * q1
    rng = np.random.default_rng(SEED_BASE + 4002)
    xs = [int(v) for v in rng.integers(-(qq // 2) + 1, qq // 2 + 1, size=P.n, dtype=np.int64)]
    xs[0], xs[1], xs[2] = qq // 2, -(qq // 2) + 1, 0
    src = np.empty((2, P.n), dtype=np.uint32)
    for t, (qi, qo) in enumerate(((q0, q1), (q1, q0))):
        inv = pow(qo, -1, qi)  # (q0 q1 / q_i)^-1 mod q_i
        src[t] = [(x * inv) % qi for x in xs]
    dst = P.np_primes(*P.level_np(P.max_level), 3)
    out = P.modup_to_max(src, base, dst)
    for t, pi in enumerate(dst[:: 7]):
        q = P.prime(pi)
        row = dst.index(pi)
        assert [int(v) for v in out[row, :64]] == [x % q for x in xs[:64]]


def test_modup_to_max_single_prime(oracle_rns24):
    P = oracle_rns24
    base = P.np_primes(1, 0)
    src = P.uniform(base, SEED_BASE + 4003)
    dst = P.np_primes(24, 0, 2)
    out = P.modup_to_max(src, base, dst)
    for t, pi in enumerate(dst):
        assert np.array_equal(out[t], src[0] % P.prime(pi))


def test_const_ops_are_inverse(oracle_rns24):
    P = oracle_rns24
    pidx = P.np_primes(24, 0, 8)
    x = P.uniform(pidx, SEED_BASE + 4004)
    rng = np.random.default_rng(SEED_BASE + 4005)
    c = np.array([rng.integers(0, P.prime(i)) for i in pidx], dtype=np.uint32)
    added = P.add_const(x, c, pidx, 0)
    assert np.array_equal(P.add_const(added, c, pidx, 1), x)
    assert np.array_equal(P.add_const(x, c, pidx, 2), P.neg(P.add_const(x, c, pidx, 1), pidx))
    zero = np.zeros(len(pidx), dtype=np.uint32)
    assert np.array_equal(P.add_const(x, zero, pidx, 0), x)
