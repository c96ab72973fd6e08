// Arithmetic of the exact int8 matrix-core forms (ck_i8.hip), shared by the
// kernels and the host table builders; tests/cpp/test_i8_arith.cpp compiles
// it for the host and checks it against 128-bit integer arithmetic.
//
// A value x with -2155905152 <= x <= 2139062143 (that is
// -128 (1 + 2^8 + 2^16 + 2^24) .. 127 (1 + 2^8 + 2^16 + 2^24), which holds
// every |x| < 2^31 - 2^24) has exactly one balanced base-256 form
// x = sum_i d_i 2^(8 i) with d_i in [-128, 127]. Its digits are the bytes
// of u = x + 0x80808080 (no byte carries, since every byte of u is
// d_i + 128 in [0, 255]) minus 128, i.e. the bytes of u with the top bit
// flipped, read as int8.
#pragma once

#include <stdint.h>

#include <vector>

#ifndef __HIPCC__
#define CK_HD inline
#else
#include <hip/hip_runtime.h>
#define CK_HD __host__ __device__ __forceinline__
#endif

namespace ck {
namespace i8 {

// the four balanced digits of a centred x (range above), as int8 bytes
CK_HD int pack_c(int x) { return (int)(((uint32_t)x + 0x80808080u) ^ 0x80808080u); }

// the digits of x - 2^30 for an unsigned x < 2^31 (x - 2^30 is in
// [-2^30, 2^30), inside the range above)
CK_HD int pack_off(uint32_t x) { return (int)((x + 0x40808080u) ^ 0x80808080u); }

// sum_j 2^(8 j) d_j from four int32 partial sums. Bounds: each D_j is a
// sum of at most 96 products of two int8 digits (|product| <= 2^14), so
// |d_1|, |d_2|, |d_3| <= 96 2^14 < 2^21, and |d_0| < 2^21 + 2^30 (an offset
// correction below 2^30 may start the d_0 accumulation; it is only used
// with 64 products). Then |lo| = |d_0 + 2^8 d_1| < 2^30 + 2^21 + 2^29
// < 2^31 and |hi| = |d_2 + 2^8 d_3| < 2^21 + 2^29, so neither int32 sum
// wraps, and hi 2^16 + lo (< 2^46 in magnitude) is an exact FP64 value:
// the fma rounds nothing.
CK_HD double combine(int d0, int d1, int d2, int d3) {
  const int lo = d0 + d1 * 256;
  const int hi = d2 + d3 * 256;
  return __builtin_fma((double)hi, 65536.0, (double)lo);
}

// |v| < 2^50 an integer-valued double, q < 2^31, qr = fl(1/q): v mod q as
// an exact integer-valued double within [-q/2 - 1, q/2 + 1]: k = rint(v qr)
// differs from v / q by at most 1/2 + |v / q| 2^-52 (< 1/2 + 2^-21 here),
// and v - k q is an integer below 2^31 in magnitude, which the fma
// produces exactly.
CK_HD double reduce_q(double v, double q, double qr) {
  const double k = __builtin_rint(v * qr);
  return __builtin_fma(-k, q, v);
}

// ---- host side: operand tables --------------------------------------------

inline uint32_t powmod_h(uint32_t a, uint64_t e, uint32_t q) {
  uint64_t r = 1, b = a % q;
  while (e) {
    if (e & 1) r = r * b % q;
    b = b * b % q;
    e >>= 1;
  }
  return (uint32_t)r;
}

inline int32_t centre_h(uint64_t x, uint32_t q) {
  x %= q;
  return x > q / 2 ? (int32_t)((int64_t)x - (int64_t)q) : (int32_t)x;
}

// byte j of pack_c(x), as it is stored in an operand register
inline uint32_t digit_byte(int32_t x, int j) {
  const uint32_t u = (uint32_t)x + 0x80808080u;
  return ((u >> (8 * j)) & 0xFFu) ^ 0x80u;
}

// Operand registers of a 16 x 16 matrix T over Z_q (rows m = outputs,
// columns = inputs) for v_mfma_i32_16x16x64_i8, as either operand: lane
// l = 16 g + m, register t, byte i of plane j is digit j of
// T[m][4 g + t] 2^(8 i) mod q (centred), so that plane j's product with the
// packed inputs is D_j[m] = sum_{n, i} T^(i)_j[m][n] d_i(x_n).
// With strided = true register t of lane group g holds input g + 4 t
// instead (for loads where that order is the contiguous one).
// out: 4 planes x 64 lanes x 4 registers.
inline void ctx_planes(const uint32_t T[16][16], uint32_t q, int32_t *out,
                       bool strided = false) {
  uint32_t p8[4];
  for (int i = 0; i < 4; i++) p8[i] = powmod_h(2, 8 * i, q);
  for (int j = 0; j < 4; j++)
    for (int lane = 0; lane < 64; lane++)
      for (int t = 0; t < 4; t++) {
        const int m = lane & 15;
        const int nn = strided ? (lane >> 4) + 4 * t : 4 * (lane >> 4) + t;
        uint32_t word = 0;
        for (int i = 0; i < 4; i++) {
          const int32_t x = centre_h((uint64_t)T[m][nn] * p8[i], q);
          word |= digit_byte(x, j) << (8 * i);
        }
        out[(j * 64 + lane) * 4 + t] = (int32_t)word;
      }
}

// Operand registers of a BaseConv table for v_mfma_i32_16x16x32_i8 (rows =
// destinations, K = 8 sources x 4 digits per k-step): tab[d * S + i] is the
// plain residue of destination d (modulus dst_q[d]) for source i. Entry
// ((mt * nks + ks) * 4 + j) * 64 + l, l = 16 g + m, holds in byte 4 e + i'
// digit j of tab[d][s] 2^(8 i') mod t_d (centred) for destination
// d = 16 mt + m and source s = 8 ks + 2 g + e (0 past D or S); nks =
// ceil(S / 8) k-steps, ceil(D / 16) row tiles mt.
inline std::vector<int64_t> bconv_i8_planes(const std::vector<uint32_t> &tab, int S, int D,
                                            const std::vector<uint32_t> &dst_q) {
  const int nks = (S + 7) / 8, nmt = (D + 15) / 16;
  std::vector<int64_t> out((size_t)nmt * nks * 4 * 64, 0);
  for (int mt = 0; mt < nmt; mt++)
    for (int ks = 0; ks < nks; ks++)
      for (int j = 0; j < 4; j++)
        for (int lane = 0; lane < 64; lane++) {
          const int d = mt * 16 + (lane & 15), gg = lane >> 4;
          uint64_t word = 0;
          for (int e = 0; e < 2; e++) {
            const int i = 8 * ks + 2 * gg + e;
            if (d >= D || i >= S) continue;
            const uint32_t q = dst_q[d];
            for (int id = 0; id < 4; id++) {
              const int32_t x =
                  centre_h((uint64_t)tab[(size_t)d * S + i] * powmod_h(2, 8 * id, q), q);
              word |= (uint64_t)digit_byte(x, j) << (8 * (4 * e + id));
            }
          }
          out[(((size_t)mt * nks + ks) * 4 + j) * 64 + lane] = (int64_t)word;
        }
  return out;
}

}  // namespace i8
}  // namespace ck

#undef CK_HD
