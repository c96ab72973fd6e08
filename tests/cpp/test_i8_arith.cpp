// CPU check of the exact int8 matrix-core arithmetic
// (cheddar-fhe_amd/csrc/ck_i8_arith.h) against 128-bit integer arithmetic.
//
// The header is compiled for the host with IEEE doubles, no contraction and
// libm's exact fma. The matrix cores are emulated with the lane layout of
// v_mfma_i32_16x16x64_i8 / v_mfma_i32_16x16x32_i8 measured on gfx950 (lane
// l = 16 g + r supplies row r of A and column r of B for k = K g + byte, and
// accumulator v of lane l is D[4 g + v][r]), with the int32 accumulation
// done in 64 bits and checked never to leave the int32 range. Checked:
//   * pack_c / pack_off: the four digits recombine to x (resp. x - 2^30)
//     over the whole stated range and its end points;
//   * combine: exact over the stated bounds, including their extremes;
//   * reduce_q: an integer congruent to v with |r| <= q/2 + 1 for |v| < 2^50;
//   * one 16-point matrix stage built with ctx_planes, on random and extreme
//     matrices and inputs (centred, and canonical with the offset
//     correction), equals sum_n T[m][n] x_n mod q for six moduli;
//   * the BaseConv planes (bconv_i8_planes) for 1..24 sources and 1..64
//     destinations with distinct moduli, through the 16x16x32 emulation;
//   * the worst case: every digit -128 in both operands (|D_j| = 2^20).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iterator>
#include <vector>

#include "ck_i8_arith.h"

using namespace ck::i8;

static int failures = 0;
static void fail(const char *what, long long a, long long b) {
  if (failures++ < 20) std::printf("FAIL %s: %lld %lld\n", what, a, b);
}

static uint64_t rng_state = 0x4348454444415208ull;
static uint64_t next_u64() {  // SplitMix64
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static int64_t uniform(int64_t lo, int64_t hi) {  // [lo, hi]
  return lo + (int64_t)(next_u64() % (uint64_t)(hi - lo + 1));
}

static int64_t mod_q(__int128 v, uint32_t q) {
  __int128 r = v % (__int128)q;
  if (r < 0) r += q;
  return (int64_t)r;
}

static int64_t digits_value(int packed) {
  int64_t v = 0;
  for (int i = 3; i >= 0; i--) v = v * 256 + (int8_t)(((uint32_t)packed >> (8 * i)) & 0xFF);
  return v;
}

static const int64_t kLo = -2155905152LL, kHi = 2139062143LL;

static void test_pack() {
  std::vector<int64_t> xs = {kLo, kLo + 1, kHi, kHi - 1, 0, 1, -1, 127, 128, -128, -129,
                             (1LL << 30), -(1LL << 30), (1LL << 31) - (1LL << 24) - 1,
                             -(1LL << 31) + (1LL << 24) + 1};
  for (int i = 0; i < 200000; i++) xs.push_back(uniform(kLo, kHi));
  for (int64_t x : xs)
    if (digits_value(pack_c((int)x)) != x) fail("pack_c", x, digits_value(pack_c((int)x)));
  std::vector<int64_t> us = {0, 1, (1LL << 31) - 1, (1LL << 30), (1LL << 30) - 1, 2147352576};
  for (int i = 0; i < 200000; i++) us.push_back(uniform(0, (1LL << 31) - 1));
  for (int64_t u : us)
    if (digits_value(pack_off((uint32_t)u)) != u - (1LL << 30))
      fail("pack_off", u, digits_value(pack_off((uint32_t)u)));
}

static void test_combine() {
  const int64_t b = 1 << 20, b0 = (1LL << 30) + b;
  const int64_t ext0[] = {-b0, b0, 0, 1, -1}, ext[] = {-b, b, 0, 1, -1};
  auto check = [](int64_t d0, int64_t d1, int64_t d2, int64_t d3) {
    const double v = combine((int)d0, (int)d1, (int)d2, (int)d3);
    const int64_t want = d0 + 256 * d1 + 65536 * d2 + 16777216 * d3;
    if (v != (double)want || (int64_t)v != want) fail("combine", want, (long long)v);
  };
  for (int64_t d0 : ext0)
    for (int64_t d1 : ext)
      for (int64_t d2 : ext)
        for (int64_t d3 : ext) check(d0, d1, d2, d3);
  for (int i = 0; i < 200000; i++)
    check(uniform(-b0, b0), uniform(-b, b), uniform(-b, b), uniform(-b, b));
}

static void test_reduce(uint32_t q) {
  const double qd = q, qr = 1.0 / qd;
  auto check = [&](int64_t v) {
    const double r = reduce_q((double)v, qd, qr);
    if (r != std::rint(r) || std::fabs(r) > qd / 2 + 1 || mod_q((__int128)v - (int64_t)r, q))
      fail("reduce_q", v, (long long)r);
  };
  const int64_t lim = (1LL << 50) - 1;
  for (int64_t v : {lim, -lim, (int64_t)0, (int64_t)q, -(int64_t)q, (int64_t)(q / 2),
                    (int64_t)(q / 2 + 1), -(int64_t)(q / 2) - 1})
    check(v);
  for (int i = 0; i < 100000; i++) {
    check(uniform(-lim, lim));
    const int64_t k = uniform(-(1LL << 18), 1LL << 18);
    check(k * (int64_t)q + uniform(-2, 2) + (int64_t)(q / 2) * (i & 1));
  }
}

// D_j[m][c] of one v_mfma_i32_16x16x64_i8 with the plane as A (rows m) and
// packed inputs as B: B lane 16 g + c register t = input 4 g + t of column c
// (input g + 4 t when strided)
static void mfma64(const int32_t *plane, const int32_t Bp[16][16], int64_t D[16][16],
                   const int64_t *init, bool strided = false) {
  for (int m = 0; m < 16; m++)
    for (int c = 0; c < 16; c++) {
      int64_t acc = init ? init[m] : 0;
      for (int g = 0; g < 4; g++)
        for (int byte = 0; byte < 16; byte++) {
          const int t = byte / 4, i = byte % 4;
          const int8_t a = (int8_t)(((uint32_t)plane[(16 * g + m) * 4 + t] >> (8 * i)) & 0xFF);
          const int in = strided ? g + 4 * t : 4 * g + t;
          const int8_t bb = (int8_t)(((uint32_t)Bp[in][c] >> (8 * i)) & 0xFF);
          acc += (int64_t)a * bb;
          if (acc > INT32_MAX || acc < INT32_MIN) fail("int32 accumulation", acc, m);
        }
      D[m][c] = acc;
    }
}

// one 16-point stage: T (16 x 16), inputs X[n][c] (16 inputs x 16 columns)
static void check_stage(const uint32_t T[16][16], uint32_t q, const int64_t X[16][16],
                        bool offset, bool strided = false) {
  std::vector<int32_t> planes(4 * 256);
  ctx_planes(T, q, planes.data(), strided);
  int32_t Bp[16][16];
  for (int n = 0; n < 16; n++)
    for (int c = 0; c < 16; c++)
      Bp[n][c] = offset ? pack_off((uint32_t)X[n][c]) : pack_c((int)X[n][c]);
  int64_t corr[16];
  for (int m = 0; m < 16; m++) {
    uint64_t s = 0;
    for (int n = 0; n < 16; n++) s += T[m][n];
    corr[m] = centre_h((uint64_t)(s % q) * ((1ull << 30) % q), q);
  }
  int64_t D[4][16][16];
  for (int j = 0; j < 4; j++)
    mfma64(planes.data() + j * 256, Bp, D[j], j == 0 && offset ? corr : nullptr, strided);
  for (int m = 0; m < 16; m++)
    for (int c = 0; c < 16; c++) {
      for (int j = 1; j < 4; j++)
        if ((D[j][m][c] < 0 ? -D[j][m][c] : D[j][m][c]) > (1 << 20)) fail("plane bound", D[j][m][c], j);
      const double r = reduce_q(combine((int)D[0][m][c], (int)D[1][m][c], (int)D[2][m][c],
                                        (int)D[3][m][c]),
                                (double)q, 1.0 / (double)q);
      __int128 want = 0;
      for (int n = 0; n < 16; n++) want += (__int128)T[m][n] * X[n][c];
      if (std::fabs(r) > q / 2.0 + 1 || mod_q(want - (int64_t)r, q)) fail("stage", m, c);
    }
}

static void test_stages(uint32_t q) {
  uint32_t T[16][16];
  int64_t X[16][16];
  const int64_t h = (q - 1) / 2;
  for (int rep = 0; rep < 60; rep++) {
    const int kind = rep % 6;  // matrix: random, all q-1, all (q-1)/2, all 1, mixed extremes
    for (int m = 0; m < 16; m++)
      for (int n = 0; n < 16; n++) {
        const uint32_t ex[] = {0, 1, q - 1, (uint32_t)h, (uint32_t)h + 1};
        T[m][n] = kind == 0 ? (uint32_t)uniform(0, q - 1)
                  : kind == 1 ? q - 1
                  : kind == 2 ? (uint32_t)h
                  : kind == 3 ? (uint32_t)h + 1
                  : kind == 4 ? 1
                              : ex[uniform(0, 4)];
      }
    // centred inputs (lazy range of the intermediate values, q/2 + small)
    for (int n = 0; n < 16; n++)
      for (int c = 0; c < 16; c++) {
        const int64_t ex[] = {-h - 1, h + 1, 0, -h, h};
        X[n][c] = (rep & 1) ? ex[uniform(0, 4)] : uniform(-h - (1 << 10), h + (1 << 10));
      }
    check_stage(T, q, X, false);
    check_stage(T, q, X, false, true);
    // canonical inputs through the offset form
    for (int n = 0; n < 16; n++)
      for (int c = 0; c < 16; c++) {
        const int64_t ex[] = {0, (int64_t)q - 1, h, h + 1, 1};
        X[n][c] = (rep & 1) ? ex[uniform(0, 4)] : uniform(0, q - 1);
      }
    check_stage(T, q, X, true);
  }
}

// BaseConv planes through v_mfma_i32_16x16x32_i8 k-steps (8 bytes a lane)
static void test_bconv(int S, int D, const std::vector<uint32_t> &mods) {
  std::vector<uint32_t> tab((size_t)D * S), dq(D);
  for (int d = 0; d < D; d++) {
    dq[d] = mods[d % mods.size()];
    for (int i = 0; i < S; i++) tab[(size_t)d * S + i] = (uint32_t)uniform(0, dq[d] - 1);
  }
  const std::vector<int64_t> pl = bconv_i8_planes(tab, S, D, dq);
  const int nks = (S + 7) / 8;
  for (int pos = 0; pos < 16; pos++) {
    int64_t v[24];
    for (int i = 0; i < 24; i++) {
      const int64_t h = (1LL << 30) - 1;
      v[i] = i < S ? (pos < 2 ? (pos ? -h : h) : uniform(-h, h)) : 0;
    }
    for (int d = 0; d < D; d++) {
      const int mt = d / 16, m = d % 16;
      int64_t Dj[4];
      for (int j = 0; j < 4; j++) {
        int64_t acc = 0;
        for (int ks = 0; ks < nks; ks++)
          for (int g = 0; g < 4; g++)
            for (int byte = 0; byte < 8; byte++) {
              const int e = byte / 4, i = byte % 4, s = 8 * ks + 2 * g + e;
              const uint64_t w = (uint64_t)pl[(((size_t)mt * nks + ks) * 4 + j) * 64 + 16 * g + m];
              const int8_t a = (int8_t)((w >> (8 * byte)) & 0xFF);
              const int packed = pack_c((int)(s < S ? v[s] : 0));
              const int8_t bb = (int8_t)(((uint32_t)packed >> (8 * i)) & 0xFF);
              acc += (int64_t)a * bb;
              if (acc > INT32_MAX || acc < INT32_MIN) fail("bconv int32", acc, d);
            }
        Dj[j] = acc;
      }
      const double r = reduce_q(combine((int)Dj[0], (int)Dj[1], (int)Dj[2], (int)Dj[3]),
                                (double)dq[d], 1.0 / (double)dq[d]);
      __int128 want = 0;
      for (int i = 0; i < S; i++) want += (__int128)tab[(size_t)d * S + i] * v[i];
      if (std::fabs(r) > dq[d] / 2.0 + 1 || mod_q(want - (int64_t)r, dq[d]))
        fail("bconv", S, d);
    }
  }
}

static void test_worst_case() {
  // every byte -128 in both operands: |D_j| = 64 * 2^14 = 2^20 exactly
  int32_t plane[256];
  int32_t Bp[16][16];
  for (int i = 0; i < 256; i++) plane[i] = (int32_t)0x80808080u;
  for (int n = 0; n < 16; n++)
    for (int c = 0; c < 16; c++) Bp[n][c] = (int32_t)0x80808080u;
  int64_t D[16][16];
  mfma64(plane, Bp, D, nullptr);
  if (D[0][0] != (1 << 20)) fail("worst case", D[0][0], 1 << 20);
  const int64_t init[16] = {(1LL << 30) - 1};
  mfma64(plane, Bp, D, init);  // with the largest offset correction
  const double v = combine((int)D[0][0], 1 << 20, 1 << 20, 1 << 20);
  const int64_t want = D[0][0] + 256LL * (1 << 20) + 65536LL * (1 << 20) + 16777216LL * (1 << 20);
  if (v != (double)want) fail("worst case combine", want, (long long)v);
}

int main() {
  // chain prime, bootparam_40 main and terminal primes, 2^31 - 1, small
  const uint32_t primes[] = {2147352577u, 2100953089u, 1196556289u, 30539777u,
                             2147483647u, 786433u};
  test_pack();
  test_combine();
  for (uint32_t q : primes) {
    test_reduce(q);
    test_stages(q);
  }
  const std::vector<uint32_t> mods(std::begin(primes), std::end(primes));
  for (int S : {1, 2, 7, 8, 9, 12, 16, 17, 24})
    for (int D : {1, 16, 17, 32, 64}) test_bconv(S, D, mods);
  test_worst_case();
  std::printf("OK: %d failure(s)\n", failures);
  return failures ? 1 : 0;
}
