"""Helpers shared by the GPU parity tests (inputs from the oracle's seeded
RNG, uploads, key generation with uniform residues)."""
import numpy as np


def uniform(P, pidx, seed):
    return P.uniform(list(pidx), seed)


def full_key(P, seed):
    """A uniform evaluation key in the reference layout: dnum blocks of
    (L + alpha) limbs for b and for a."""
    allp = list(range(P.num_primes))
    kb = np.stack([P.uniform(allp, seed + 2 * g) for g in range(P.dnum)])
    ka = np.stack([P.uniform(allp, seed + 2 * g + 1) for g in range(P.dnum)])
    return kb, ka


def dev_key(kb, ka):
    import cheddar_amd as ck
    return [ck.to_device(kb[g]) for g in range(kb.shape[0])], \
           [ck.to_device(ka[g]) for g in range(ka.shape[0])]


def assert_same(got, want, what):
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, (what, got.shape, want.shape)
    if not np.array_equal(got, want):
        bad = np.argwhere(got != want)
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} mismatches, first at {i}: "
                             f"got {got[i]} want {want[i]}")
