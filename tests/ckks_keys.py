"""Test-side RLWE key generation, encryption and decryption on top of the
oracle's primitives, restating the reference's test harness conventions
(src/UserInterface.cu:366-458 PrepareEvk / AddEvkPart, :279-307 rotation
keys, Encrypt/Decrypt kernels :16-108, :184-189; insecure by design, as the
reference says of its own harness). Everything is in the NTT domain in
Montgomery form, limb-major in global prime order.
"""
import numpy as np

import ckks_oracle as O


class Keys:
    def __init__(self, P, seed, hamming=None):
        self.P = P
        rng = np.random.default_rng(seed)
        n = P.n
        if hamming is None:
            s = rng.integers(-1, 2, size=n)
        else:
            s = np.zeros(n, dtype=np.int64)
            idx = rng.choice(n, size=hamming, replace=False)
            s[idx] = rng.choice([-1, 1], size=hamming)
        self.s_coef = s.astype(np.int64)
        self.rng = rng
        self.allp = list(range(P.num_primes))
        self.s = self.ntt_small(self.s_coef, self.allp)

    # ---- helpers -------------------------------------------------------
    def qcol(self, pidx):
        return np.array([self.P.prime(i) for i in pidx], dtype=np.int64)[:, None]

    def ntt_small(self, coef, pidx):
        """small integer polynomial -> NTT domain, Montgomery form"""
        q = self.qcol(pidx)
        res = (np.asarray(coef, dtype=np.int64)[None, :] % q).astype(np.uint32)
        return self.P.ntt(res, pidx, mont_conv=True)

    def small(self, bound):
        return self.rng.integers(-bound, bound + 1, size=self.P.n)

    def uniform(self, pidx):
        return self.P.uniform(pidx, int(self.rng.integers(1, 2 ** 62)))

    def sub_s(self, pidx):
        return self.s[pidx]

    # ---- encryption ------------------------------------------------------
    def encrypt(self, m_coef, level, err=3):
        P = self.P
        pidx = P.level_primes(level)
        a = self.uniform(pidx)
        e = self.ntt_small(self.small(err), pidx)
        m = self.ntt_small(m_coef, pidx)
        as_ = P.mult(a, self.s[pidx], pidx)
        b = P.add(P.sub(m, as_, pidx), e, pidx)
        return b, a

    def decrypt(self, ct, level, positions):
        """centred integer coefficients (big ints) at the given positions"""
        P = self.P
        nm, nt = P.level_np(level)
        pidx = P.level_primes(level)
        b, a = ct
        m = P.add(b, P.mult(a, self.s[pidx], pidx), pidx)
        mc = P.intt_np(m, nm, nt, 0, mont_conv=True)
        primes = [P.prime(i) for i in pidx]
        M = 1
        for q in primes:
            M *= q
        out = []
        for k in positions:
            x = 0
            for t, q in enumerate(primes):
                Mi = M // q
                x += int(mc[t, k]) * Mi * pow(Mi, -1, q)
            x %= M
            out.append(x - M if x > M // 2 else x)
        return out

    # ---- evaluation keys ---------------------------------------------------
    def evk(self, target, enc_secret=None, err=3):
        """EvaluationKey switching enc_secret -> target:
        b_g = -a_g s_enc + e_g + P * target on group g's q limbs"""
        P = self.P
        s_enc = self.s if enc_secret is None else enc_secret
        kb, ka = [], []
        pp = np.zeros(P.num_primes, dtype=np.uint32)
        for i in range(P.L):
            pp[i] = P.p_prod_mont(i)
        pt = P.mult_const(target, pp, self.allp)
        for g in range(P.dnum):
            a = self.uniform(self.allp)
            e = self.ntt_small(self.small(err), self.allp)
            b = P.add(P.sub(e, P.mult(a, s_enc, self.allp), self.allp),
                      np.zeros_like(a), self.allp)
            lo, hi = g * P.alpha, min((g + 1) * P.alpha, P.L)
            part = [i for i in range(lo, hi)]
            b[part] = P.add(b[part], pt[part], part)
            kb.append(b)
            ka.append(a)
        return np.stack(kb), np.stack(ka)

    def mult_key(self):
        s2 = self.P.mult(self.s, self.s, self.allp)
        return self.evk(s2)

    def rot_key(self, rot):
        """Context::HRot key for 5^rot: encryption secret sigma_{5^(N/2-rot)}(s)
        (UserInterface.cu:279-307)"""
        g = self.P.galois(self.P.n // 2 - rot)
        return self.evk(self.s, enc_secret=self.P.permute(self.s, g))

    def conj_key(self):
        return self.evk(self.s, enc_secret=self.P.permute(self.s, self.P.galois(-1)))


def negacyclic_product(a, b):
    """exact integer negacyclic product of two integer polynomials"""
    n = len(a)
    # inputs |v| < 2^31 are shifted to [0, 2^33) so every slot of the packed
    # product is a non-negative sum of at most n terms < 2^66 (< 2^128)
    off = 1 << 32
    width = 16
    A = int.from_bytes(b"".join((int(v) + off).to_bytes(width, "little") for v in a), "little")
    B = int.from_bytes(b"".join((int(v) + off).to_bytes(width, "little") for v in b), "little")
    # slot k holds sum (a_i + off)(b_j + off) over i + j = k: remove the
    # offset terms exactly with prefix sums
    C = A * B
    raw = C.to_bytes(2 * n * width + width, "little")
    full = [int.from_bytes(raw[i * width:(i + 1) * width], "little") for i in range(2 * n - 1)]
    sa = [0] * (2 * n)
    # prefix sums for the offset corrections
    pa = [0]
    pb = [0]
    for v in a:
        pa.append(pa[-1] + int(v))
    for v in b:
        pb.append(pb[-1] + int(v))
    out = []
    for k in range(2 * n - 1):
        lo, hi = max(0, k - n + 1), min(k, n - 1)
        cnt = hi - lo + 1
        sum_a = pa[hi + 1] - pa[lo]
        sum_b = pb[k - lo + 1] - pb[k - hi]
        sa[k] = full[k] - off * sum_a - off * sum_b - off * off * cnt
    for k in range(n):
        out.append(sa[k] - (sa[k + n] if k + n < 2 * n - 1 else 0))
    return out


def automorph_coef(m, g):
    """m(X^g) mod X^N + 1 on integer coefficients"""
    n = len(m)
    out = [0] * n
    for j, v in enumerate(m):
        e = (j * g) % (2 * n)
        if e >= n:
            out[e - n] -= int(v)
        else:
            out[e] += int(v)
    return out
