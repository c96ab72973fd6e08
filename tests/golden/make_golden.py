"""Generate tests/golden/vectors.json: known-answer vectors computed with
Python's built-in big integers, independently of the C oracle.

The reference ships no golden vectors (its tests compare decrypted values
within 1e-3), so these vectors pin the oracle to the mathematics its
functions restate:
  * generators / roots: the smallest generator of Z_q^* by brute force and
    psi = g^((q-1)/2N) for every prime of the parameter files;
  * Montgomery: a * 2^-32 mod q and the reference's signed lazy range;
  * NTT: direct evaluation out[k] = sum_j a_j psi^((2 brv(k)+1) j) at N = 16;
  * BaseConv: y_m = (sum_i v_i prod_{j!=i} s_j) * 2^32 mod t_m.

Run: python tests/golden/make_golden.py   (writes vectors.json next to it)
"""
import json
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))
PARAMS = os.path.join(HERE, "params")
PKG_PARAMS = os.path.join(os.path.dirname(os.path.dirname(HERE)), "cheddar-fhe_amd", "params")


def factor(n):
    f, d = [], 2
    while d * d <= n:
        if n % d == 0:
            f.append(d)
            while n % d == 0:
                n //= d
        d += 1
    if n > 1:
        f.append(n)
    return f


def smallest_generator(q):
    phi = q - 1
    fs = factor(phi)
    g = 2
    while True:
        if all(pow(g, phi // f, q) != 1 for f in fs):
            return g
        g += 1


def brv(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2)


def all_primes():
    out = []
    for d in (PARAMS, PKG_PARAMS):
        for name in sorted(os.listdir(d)):
            with open(os.path.join(d, name)) as f:
                p = json.load(f)
            for key in ("main_primes", "terminal_primes", "auxiliary_primes"):
                for q in p.get(key, []):
                    if q < 2 ** 31 and q not in out:
                        out.append(q)
    return out


def main():
    rng = random.Random(0x4348454444415200)
    gold = {}
    # generators and 2N-th roots for N = 2^16
    roots = []
    for q in all_primes():
        g = smallest_generator(q)
        psi = pow(g, (q - 1) // (2 ** 17), q)
        assert pow(psi, 2 ** 16, q) == q - 1
        roots.append({"q": q, "g": g, "psi": psi, "qinv": pow(q, -1, 2 ** 32)})
    gold["roots"] = roots

    # Montgomery reduction: a * 2^-32 mod q for signed |a| <= q * 2^31
    mont = []
    for q in (2147352577, 2100953089, 974258177, 30539777):
        rinv = pow(2 ** 32, -1, q)
        cases = [0, 1, -1, q - 1, -(q - 1), q * 2 ** 31 - 1, -q * 2 ** 31,
                 (q - 1) ** 2, -((q - 1) // 2) ** 2]
        cases += [rng.randrange(-q * 2 ** 31, q * 2 ** 31) for _ in range(12)]
        for a in cases:
            mont.append({"q": q, "a": a, "r": (a * rinv) % q})
    gold["mont"] = mont

    # NTT at N = 16 by direct evaluation
    ntt = []
    n, logn = 16, 4
    for q in (2147352577, 2113011713, 30539777, 974258177):
        g = smallest_generator(q)
        psi = pow(g, (q - 1) // (2 * n), q)
        a = [rng.randrange(q) for _ in range(n)]
        out = [sum(a[j] * pow(psi, (2 * brv(k, logn) + 1) * j, q) for j in range(n)) % q
               for k in range(n)]
        ntt.append({"q": q, "log_n": logn, "in": a, "out": out})
    gold["ntt16"] = ntt

    # BaseConv: 3 source primes -> 2 destination primes
    src = [2147352577, 2146959361, 2146041857]
    dst = [2144468993, 2142502913]
    bconv = []
    for _ in range(4):
        v = [rng.randrange(-(s - 1) // 2, (s - 1) // 2 + 1) for s in src]
        X = sum(vi * (src[0] * src[1] * src[2] // si) for vi, si in zip(v, src))
        bconv.append({"src": src, "dst": dst, "v": v,
                      "y": [(X * 2 ** 32) % t for t in dst]})
    gold["bconv"] = bconv

    with open(os.path.join(HERE, "vectors.json"), "w") as f:
        json.dump(gold, f, indent=1)
    print(f"wrote {len(roots)} roots, {len(mont)} montgomery, {len(ntt)} ntt, "
          f"{len(bconv)} bconv vectors")


if __name__ == "__main__":
    main()
