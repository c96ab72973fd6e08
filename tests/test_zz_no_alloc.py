"""ck_hmult_noalloc (HMult with its temporaries in the library's static
device array, no allocation) against ck_hmult through the C ABI: two calls
in a row on one device, the second at a lower level (a smaller input), each
word for word equal to ck_hmult's output; levels whose workspace exceeds the
array are rejected."""
import numpy as np
import pytest

from conftest import SEED_BASE
from gpu_util import assert_same, dev_key, full_key, uniform

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_hmult_noalloc_matches_hmult(rns24, oracle_rns24):
    import cheddar_amd as ck
    ctx = ck.Context.from_json(rns24)
    P = oracle_rns24
    key = dev_key(*full_key(P, SEED_BASE + 5000))
    for call, level in enumerate([23, 9]):
        pidx = P.level_primes(level)
        ct1 = [ck.to_device(uniform(P, pidx, SEED_BASE + 5100 + 10 * call + i)) for i in range(2)]
        ct2 = [ck.to_device(uniform(P, pidx, SEED_BASE + 5105 + 10 * call + i)) for i in range(2)]
        shape = (P.num_q(level - 1), P.n)
        want = [torch.full(shape, -1, dtype=torch.int32, device="cuda") for _ in range(2)]
        got = [torch.full(shape, -1, dtype=torch.int32, device="cuda") for _ in range(2)]
        ctx.hmult(want, ct1, ct2, key, level, rescale=True)
        ctx.hmult_noalloc(got, ct1, ct2, key, level, rescale=True)
        torch.cuda.synchronize()
        for j in range(2):
            assert_same(ck.to_host(got[j]), ck.to_host(want[j]), f"call {call} poly {j}")
        # and the oracle agrees (the reference's HMult)
        want_o = P.hmult(level, [ck.to_host(t) for t in ct1], [ck.to_host(t) for t in ct2],
                         (np.stack([ck.to_host(k) for k in key[0]]),
                          np.stack([ck.to_host(k) for k in key[1]])), rescale=True)
        for j in range(2):
            assert_same(ck.to_host(got[j]), want_o[j], f"call {call} poly {j} vs oracle")


def test_hmult_noalloc_rejects_oversized(boot40):
    """bootparam_40 at its top level needs more than the 192-limb array"""
    import cheddar_amd as ck
    ctx = ck.Context.from_json(boot40)
    top = len(boot40["level_config"]) - 1
    t = torch.zeros(16, dtype=torch.int32, device="cuda")
    with pytest.raises(ck.CheddarError, match="static scratch"):
        ctx.hmult_noalloc([t, t], [t, t], [t, t], ([t] * 4, [t] * 4), top)
