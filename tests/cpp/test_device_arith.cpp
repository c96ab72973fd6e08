// CPU check of the device modular arithmetic of cheddar-fhe_amd/csrc/ck_arith.h.
// The header is compiled for the host (the HIP runtime header is skipped and
// __mulhi is supplied) with IEEE doubles, no contraction and libm's exact fma,
// i.e. the arithmetic the gfx950 build executes, and every function is
// compared with exact 128-bit integer arithmetic on random operands over the
// full lazy ranges the kernels use and on edge cases (0, 1, q - 1, range
// limits, exact multiples of q and their neighbours).
#define HIP_INCLUDE_HIP_HIP_RUNTIME_H
#define __device__
#define __forceinline__ inline
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>
static inline int32_t __mulhi(int32_t a, int32_t b) {
  return (int32_t)(((int64_t)a * (int64_t)b) >> 32);
}
#include "ck_arith.h"

using namespace ck;

static int failures = 0;
static void fail(const char *what, double a, double w, uint32_t q, double got) {
  if (failures++ < 20)
    std::printf("FAIL %s: a=%.0f w=%.0f q=%u got=%.17g\n", what, a, w, q, got);
}

static uint64_t rng_state = 0x4348454444415206ull;
static uint64_t next_u64() {  // SplitMix64
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// uniform integer in [-2^bits + 1, 2^bits - 1] as a double
static double lazy(int bits) {
  const int64_t m = ((int64_t)1 << bits) - 1;
  int64_t v = (int64_t)(next_u64() & (uint64_t)m);
  return (double)((next_u64() & 1) ? -v : v);
}

static int64_t mod_q(__int128 v, uint32_t q) {
  __int128 r = v % (__int128)q;
  if (r < 0) r += q;
  return (int64_t)r;
}

static bool is_int(double r) { return std::isfinite(r) && r == std::rint(r); }

// r must be an integer congruent to a*w with |r| < 0.75 q
static void check_product(const char *what, double r, double a, double w, uint32_t q) {
  const __int128 exact = (__int128)(int64_t)a * (int64_t)w;
  if (!is_int(r) || std::fabs(r) >= 0.75 * q ||
      mod_q((__int128)(int64_t)r - exact, q) != 0)
    fail(what, a, w, q, r);
}

static void test_prime(uint32_t q) {
  const double qd = (double)q, qr = 1.0 / qd;
  // q^-1 mod 2^32 (Newton), signed as the context stores it
  uint32_t inv = q;
  for (int i = 0; i < 5; i++) inv *= 2u - q * inv;
  const int32_t qinv = (int32_t)inv;
  const ModF m{qd, qr};

  std::vector<double> as = {0.0, 1.0, -1.0, qd - 1, -(qd - 1), qd, -qd, qd + 1,
                            (double)(((int64_t)1 << 40) - 1), -(double)(((int64_t)1 << 40) - 1),
                            (double)(int64_t)(((int64_t)1 << 40) / q * q),
                            (double)(int64_t)(((int64_t)1 << 40) / q * q - 1),
                            (double)(((int64_t)1 << 32) - 1)};
  std::vector<double> ws = {0.0, 1.0, 2.0, qd - 1, qd - 2, std::floor(qd / 2)};
  // edge operands against edge and random multipliers
  for (double a : as) {
    for (double w : ws) {
      check_product("mulmod_f edge", mulmod_f(a, w, w * qr, qd), a, w, q);
      check_product("mulmod_fq edge", mulmod_fq(a, w, qd, qr), a, w, q);
    }
    for (int i = 0; i < 2000; i++) {
      const double w = (double)(next_u64() % q);
      check_product("mulmod_f edge-a", mulmod_f(a, w, w * qr, qd), a, w, q);
      check_product("mulmod_fq edge-a", mulmod_fq(a, w, qd, qr), a, w, q);
    }
  }
  // random lazy operands: |a| < 2^40 (transform values), < 2^36, < 2^32
  // (sums of two residues: the Karatsuba operand), canonical residues
  for (int bits : {40, 36, 32, 31}) {
    for (int i = 0; i < 400000; i++) {
      const double a = bits == 31 ? (double)(next_u64() % q) : lazy(bits);
      const double w = (double)(next_u64() % q);
      check_product("mulmod_f", mulmod_f(a, w, w * qr, qd), a, w, q);
      check_product("mulmod_fq", mulmod_fq(a, w, qd, qr), a, w, q);
    }
  }
  // the quotient-from-product form with a multiplier that is itself a sum of
  // two residues (below 2^32)
  for (int i = 0; i < 400000; i++) {
    const double a = (double)(next_u64() % q) + (double)(next_u64() % q);
    const double w = (double)(next_u64() % q) + (double)(next_u64() % q);
    const double r = mulmod_fq(a, w, qd, qr);
    const __int128 exact = (__int128)(int64_t)a * (int64_t)w;
    if (!is_int(r) || std::fabs(r) >= 0.75 * q || mod_q((__int128)(int64_t)r - exact, q) != 0)
      fail("mulmod_fq sum operands", a, w, q, r);
  }

  // reduce_f / canon_f / centre_f over |x| < 2^50, canon_small over (-q, q)
  const int64_t h = (q - 1) / 2;
  for (int i = 0; i < 400000; i++) {
    const double x = i < 200000 ? lazy(50) : lazy(40);
    const int64_t ex = mod_q((__int128)(int64_t)x, q);
    const double r = reduce_f(x, m);
    if (!is_int(r) || std::fabs(r) >= 0.75 * q || mod_q((__int128)(int64_t)r - (int64_t)x, q) != 0)
      fail("reduce_f", x, 0, q, r);
    if ((int64_t)canon_f(x, m) != ex) fail("canon_f", x, 0, q, canon_f(x, m));
    const int64_t c = centre_f(x, m);
    if (c < -h || c > h || mod_q((__int128)(c - ex), q) != 0) fail("centre_f", x, 0, q, (double)c);
  }
  for (int64_t r : {(int64_t)0, (int64_t)1, (int64_t)-1, (int64_t)q - 1, -((int64_t)q - 1),
                    (int64_t)h, -(int64_t)h}) {
    if ((int64_t)canon_small((double)r, qd) != mod_q(r, q)) fail("canon_small edge", (double)r, 0, q, 0);
  }
  for (int i = 0; i < 400000; i++) {
    const int64_t r = (int64_t)(next_u64() % (2ull * q - 1)) - ((int64_t)q - 1);
    if ((int64_t)canon_small((double)r, qd) != mod_q(r, q)) fail("canon_small", (double)r, 0, q, 0);
  }

  // Montgomery family: mont_reduce(a) = a 2^-32 mod q for |a| < q 2^31
  const __int128 r32inv = [&] {  // 2^-32 mod q (extended Euclid)
    int64_t r0 = q, r1 = mod_q((__int128)1 << 32, q), s0 = 0, s1 = 1;
    while (r1) {
      const int64_t t = r0 / r1, r2 = r0 - t * r1, s2 = s0 - t * s1;
      r0 = r1, r1 = r2, s0 = s1, s1 = s2;
    }
    return (__int128)mod_q(s0, q);
  }();
  // the strict bound: a = q 2^31 exactly reduces to q itself
  const int64_t lim = (int64_t)q << 31;
  if (mont_reduce_lazy(lim, q, qinv) != (int64_t)q) fail("mont_reduce_lazy bound", (double)lim, 0, q, 0);
  for (int i = 0; i < 400000; i++) {
    // |a| < q 2^31, the range mont_reduce_lazy is specified for
    int64_t a = (int64_t)(next_u64() % (uint64_t)(2 * lim - 1)) - (lim - 1);
    if (i < 4) a = (i & 1 ? 1 : -1) * (i < 2 ? lim - 1 : 0);
    const int64_t ex = mod_q((__int128)a * r32inv, q);
    const int32_t lz = mont_reduce_lazy(a, q, qinv);
    if (lz <= -(int64_t)q || lz >= (int64_t)q || mod_q((__int128)lz - ex, q) != 0)
      fail("mont_reduce_lazy", (double)a, 0, q, lz);
    if ((int64_t)mont_reduce(a, q, qinv) != ex) fail("mont_reduce", (double)a, 0, q, mont_reduce(a, q, qinv));
    const uint32_t x = (uint32_t)(next_u64() % q), y = (uint32_t)(next_u64() % q);
    if ((int64_t)mont_mul(x, y, q, qinv) != mod_q((__int128)x * y % q * r32inv, q))
      fail("mont_mul", x, y, q, mont_mul(x, y, q, qinv));
    if (add_mod(x, y, q) != (x + (uint64_t)y) % q) fail("add_mod", x, y, q, add_mod(x, y, q));
    if ((int64_t)sub_mod(x, y, q) != mod_q((int64_t)x - y, q)) fail("sub_mod", x, y, q, sub_mod(x, y, q));
    const int32_t cx = centre(x, q);
    if (cx < -h || cx > h || mod_q((__int128)cx - x, q) != 0) fail("centre", x, 0, q, cx);
    // mont_mul_pre with wq = w q^-1 mod 2^32, centred lazy inputs
    const int32_t ac = centre(x, q), wc = centre(y, q);
    const int32_t wq = (int32_t)((uint32_t)wc * inv);
    const int32_t p = mont_mul_pre(ac, wc, wq, q);
    if (p <= -(int64_t)q || p >= (int64_t)q ||
        mod_q((__int128)p - (__int128)ac * wc * r32inv, q) != 0)
      fail("mont_mul_pre", ac, wc, q, p);
  }
}

int main() {
  // the largest and smallest primes of the 32-prime chain q = 1 mod 2^17, a
  // main and a terminal prime of bootparam_40.json, 2^31 - 1 (the top of the
  // word range) and a small prime
  const uint32_t primes[] = {2147352577u, 2100953089u, 1083703297u, 30539777u,
                             2147483647u, 786433u};
  for (uint32_t q : primes) test_prime(q);
  std::printf("OK: %d failure(s)\n", failures);
  return failures ? 1 : 0;
}
