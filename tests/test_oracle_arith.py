"""Oracle pins for scalar arithmetic and per-prime constants (CPU only).

Checks the oracle's restatement of include/common/Basic.cuh:72-260 and
include/common/PrimeUtils.h:57-248 against (a) Python big-integer
known-answer vectors (tests/golden/vectors.json), (b) brute force, and (c)
the reference's own host-side PrimeUtils compiled from its headers into
oracle/_ref (skipped when that library was not built).
"""
import ctypes
import json
import os
import random

import numpy as np
import pytest

from conftest import GOLDEN, ORACLE

import ckks_oracle as O

lib = O.lib

with open(os.path.join(GOLDEN, "vectors.json")) as f:
    GOLD = json.load(f)

REF_LIB = os.path.join(ORACLE, "_ref", "libref_primeutils.so")


def test_montgomery_known_answers():
    for c in GOLD["mont"]:
        q, a = c["q"], c["a"]
        qinv = lib.or_inv_mod_base(q)
        r = lib.or_mont_reduce(a, q, qinv)
        assert r == c["r"], c
        lazy = lib.or_mont_reduce_lazy(a, q, qinv)
        assert -q < lazy < q and (lazy - c["r"]) % q == 0


def test_montgomery_brute_force_random():
    rng = random.Random(7)
    for q in (2147352577, 2100953089, 1083703297, 30539777, 485621761):
        qinv = lib.or_inv_mod_base(q)
        assert (q * qinv) % 2 ** 32 == 1
        rinv = pow(2 ** 32, -1, q)
        for _ in range(2000):
            x = rng.randrange(q)
            y = rng.randrange(q)
            assert lib.or_mult_mont(x, y, q, qinv) == (x * y * rinv) % q
            assert lib.or_mulmod(x, y, q) == (x * y) % q
            assert lib.or_to_mont(x, q) == (x << 32) % q
            n = lib.or_normalize(x, q)
            assert -(q - 1) // 2 <= n <= (q - 1) // 2 and (n - x) % q == 0


def test_invmod_powmod():
    rng = random.Random(8)
    for q in (2147352577, 974258177, 30539777):
        for _ in range(200):
            a = rng.randrange(1, q)
            assert lib.or_invmod(a, q) == pow(a, -1, q)
            e = rng.randrange(-(q - 1), q - 1)
            want = pow(a, e, q) if e >= 0 else pow(pow(a, -1, q), -e, q)
            assert lib.or_powmod(a, e, q) == want


def test_generators_and_roots_match_bruteforce():
    for c in GOLD["roots"]:
        q = c["q"]
        assert lib.or_find_generator(q) == c["g"], q
        assert lib.or_primitive_root(2 ** 17, q) == c["psi"], q
        assert lib.or_inv_mod_base(q) % 2 ** 32 == c["qinv"]


def test_bitrev():
    for bits in (1, 4, 16, 17):
        for i in (0, 1, 2, 3, (1 << bits) - 1, 12345 % (1 << bits)):
            if i >= (1 << bits):
                continue
            assert lib.or_bitrev(i, bits) == int(format(i, f"0{bits}b")[::-1], 2)


def test_param_tables(oracle_rns24):
    P = oracle_rns24
    for pi in (0, 5, 23, 24, 31):
        q = P.prime(pi)
        psi = P.psi(pi)
        assert pow(psi, 2 ** 16, q) == q - 1 and pow(psi, 2 ** 17, q) == 1
        tab = P.psi_rev_mont(pi)
        inv = P.psi_inv_rev_mont(pi)
        for j in (0, 1, 2, 3, 255, 256, 4097, 65535):
            e = int(format(j, "016b")[::-1], 2)
            assert int(tab[j]) == (pow(psi, e, q) << 32) % q
            assert int(inv[j]) == (pow(psi, -e, q) << 32) % q
    # P mod q_i in Montgomery form (Context.cpp:85-98)
    Pv = 1
    for i in range(P.L, P.num_primes):
        Pv *= P.prime(i)
    for i in range(P.L):
        q = P.prime(i)
        assert P.p_prod_mont(i) == ((Pv % q) << 32) % q


def test_prime_order_and_levels(boot40):
    """[terminal reversed | main | aux] order (Parameter.cu:103-111) and
    LevelToNP / GetPrimeVector."""
    P = O.Param.from_json(boot40)
    ter, main, aux = boot40["terminal_primes"], boot40["main_primes"], boot40["auxiliary_primes"]
    want = list(reversed(ter)) + main + aux
    assert [P.prime(i) for i in range(P.num_primes)] == want
    assert P.dnum == -(-(len(main) + len(ter)) // len(aux))
    for lv, (m, t) in enumerate(boot40["level_config"]):
        assert P.level_np(lv) == (m, t)
        primes = [P.prime(i) for i in P.level_primes(lv)]
        assert primes == list(reversed(ter[:t])) + main[:m]


def test_invalid_parameters_rejected():
    good_main = [2147352577, 2146959361]
    aux = [2146041857]
    O.Param(16, good_main, aux, level_config=[[1, 0], [2, 0]])
    with pytest.raises(ValueError):  # not = 1 mod 2N
        O.Param(16, [2147483647, 2146959361], aux, level_config=[[1, 0], [2, 0]])
    with pytest.raises(ValueError):  # last level must hold every prime
        O.Param(16, good_main, aux, level_config=[[1, 0]])
    with pytest.raises(ValueError):  # levels must grow
        O.Param(16, good_main, aux, level_config=[[2, 0], [1, 0], [2, 0]])
    # 64-bit parameter set is out of the 32-bit word range
    with open(os.path.join(GOLDEN, "params", "bootparam_40_64bit.json")) as f:
        p64 = json.load(f)
    with pytest.raises(ValueError):
        O.Param.from_json(p64)


def test_galois_factors(oracle_rns24):
    P = oracle_rns24
    assert P.galois(0) == 1
    assert P.galois(-1) == 2 * P.n - 1
    g = 1
    for r in range(1, 50):
        g = g * 5 % (2 * P.n)
        assert P.galois(r) == g


# ---------------------------------------------------------------------------
# cross-check with the reference's own PrimeUtils (oracle/_ref)
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref not built (needs the reference headers)")
    r = ctypes.CDLL(REF_LIB)
    r.ref_find_generator.restype = ctypes.c_uint32
    r.ref_find_generator.argtypes = [ctypes.c_uint32]
    r.ref_primitive_root.restype = ctypes.c_uint32
    r.ref_primitive_root.argtypes = [ctypes.c_int, ctypes.c_uint32]
    r.ref_inv_mod_base.restype = ctypes.c_int32
    r.ref_inv_mod_base.argtypes = [ctypes.c_uint32]
    r.ref_to_mont.restype = ctypes.c_uint32
    r.ref_to_mont.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    r.ref_pow_mod.restype = ctypes.c_uint32
    r.ref_pow_mod.argtypes = [ctypes.c_uint32, ctypes.c_int64, ctypes.c_uint32]
    r.ref_inv_mod.restype = ctypes.c_uint32
    r.ref_inv_mod.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    r.ref_psi_rev_mont.restype = None
    r.ref_psi_rev_mont.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32)]
    return r


def test_constants_match_reference_primeutils(ref):
    rng = random.Random(9)
    for c in GOLD["roots"]:
        q = c["q"]
        assert lib.or_find_generator(q) == ref.ref_find_generator(q)
        assert lib.or_primitive_root(2 ** 17, q) == ref.ref_primitive_root(2 ** 17, q)
        assert lib.or_inv_mod_base(q) == ref.ref_inv_mod_base(q)
        for _ in range(20):
            a = rng.randrange(1, q)
            e = rng.randrange(-(q - 2), q - 1)
            assert lib.or_to_mont(a, q) == ref.ref_to_mont(a, q)
            assert lib.or_invmod(a, q) == ref.ref_inv_mod(a, q)
            assert lib.or_powmod(a, e, q) == ref.ref_pow_mod(a, e, q)


def test_twiddle_table_matches_reference(ref, oracle_rns24, oracle_boot40):
    for P in (oracle_rns24, oracle_boot40):
        for pi in (0, P.num_primes // 2, P.num_primes - 1):
            want = np.empty(P.n, dtype=np.uint32)
            ref.ref_psi_rev_mont(P.prime(pi), 16, want.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
            assert np.array_equal(P.psi_rev_mont(pi), want)
