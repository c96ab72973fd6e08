"""GPU end-to-end semantics with real keys: ciphertexts encrypted under a
ternary secret go through the HIP HMult / HRot / HConj (C ABI), must match
the oracle bit for bit, and must decrypt to the expected plaintext (the
reference's own test semantics, unittest/BasicTest.cpp:327-419)."""
import numpy as np
import pytest

from conftest import SEED_BASE
from ckks_keys import Keys, automorph_coef, negacyclic_product
from gpu_util import assert_same, dev_key

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
def keys(oracle_rns24):
    return Keys(oracle_rns24, SEED_BASE + 600)


def test_gpu_hmult_decrypts(ck, ctx24, oracle_rns24, keys):
    P = oracle_rns24
    level = 23
    rng = np.random.default_rng(21)
    m1 = rng.integers(-2 ** 24, 2 ** 24, size=P.n)
    m2 = rng.integers(-2 ** 24, 2 ** 24, size=P.n)
    ct1, ct2 = keys.encrypt(m1, level), keys.encrypt(m2, level)
    evk = keys.mult_key()
    want = P.hmult(level, ct1, ct2, evk, rescale=True)
    out = [torch.empty((P.num_q(level - 1), P.n), dtype=torch.int32, device="cuda")
           for _ in range(2)]
    ctx24.hmult(out, [ck.to_device(v) for v in ct1], [ck.to_device(v) for v in ct2],
                dev_key(*evk), level, rescale=True)
    torch.cuda.synchronize()
    got = (ck.to_host(out[0]), ck.to_host(out[1]))
    assert_same(got[0], want[0], "hmult b")
    assert_same(got[1], want[1], "hmult a")
    prod = negacyclic_product(m1, m2)
    ql = P.prime(P.level_primes(level)[-1])
    ks = [0, 1, P.n - 1, 31337]
    for k, v in zip(ks, keys.decrypt(got, level - 1, ks)):
        assert abs(v - round(prod[k] / ql)) < 2 ** 14


@pytest.mark.parametrize("rot", [1, 1234, -1])
def test_gpu_hrot_hconj_decrypt(ck, ctx24, oracle_rns24, keys, rot):
    P = oracle_rns24
    level = 15
    rng = np.random.default_rng(30 + rot)
    m = rng.integers(-2 ** 24, 2 ** 24, size=P.n)
    ct = keys.encrypt(m, level)
    key = keys.conj_key() if rot == -1 else keys.rot_key(rot)
    d = [ck.to_device(v) for v in ct]
    out = [torch.empty_like(d[0]) for _ in range(2)]
    if rot == -1:
        ctx24.hconj(out, d, dev_key(*key), level)
        want = P.hconj(level, ct, key)
        assert_same(ck.to_host(out[0]), want[0], "hconj b")
        assert_same(ck.to_host(out[1]), want[1], "hconj a")
    else:
        ctx24.hrot(out, d, dev_key(*key), rot, level)
        want = P.hrot(level, ct, key, rot)
        assert_same(ck.to_host(out[0]), want[0], "hrot b")
        assert_same(ck.to_host(out[1]), want[1], "hrot a")
    torch.cuda.synchronize()
    got = (ck.to_host(out[0]), ck.to_host(out[1]))
    expect = automorph_coef(m, P.galois(rot))
    ks = [0, 2, P.n - 1, 4242]
    for k, v in zip(ks, keys.decrypt(got, level, ks)):
        assert abs(v - expect[k]) < 2 ** 14
