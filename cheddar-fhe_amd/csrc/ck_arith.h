// Device-side modular arithmetic for 32-bit RNS residues (q < 2^31) on
// CDNA4 (gfx950).
//
// Two exact arithmetic families are provided:
//   * signed lazy Montgomery with R = 2^32, the word-level convention the
//     reference stores its data in (include/common/Basic.cuh:72-260):
//     results of a product of stored values carry the factor R^-1;
//   * exact FP64 modular products for transforms: a residue is held as a
//     double (integers up to 2^53 are exact), a * w mod q is computed as
//     hi = a*w, lo = fma(a, w, -hi) (the exact rounding error), k =
//     rint(a * w/q), r = fma(-k, q, hi) + lo. FP64 FMA is full rate on
//     CDNA4 while 32-bit integer multiplies are quarter rate, so transforms
//     use this form. Both families produce the same canonical residues.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ck {

// ---------------------------------------------------------------------------
// Montgomery (R = 2^32), signed lazy, as Basic.cuh:90-99 defines it
// ---------------------------------------------------------------------------

// |a| < q * 2^31  ->  a * 2^-32 mod q in (-q, q): the result is
// (a - t q) / 2^32 with |t| <= 2^31 (strict bound needed: a = q 2^31 gives q)
__device__ __forceinline__ int32_t mont_reduce_lazy(int64_t a, uint32_t q,
                                                    int32_t qinv) {
  int32_t lo = (int32_t)(uint32_t)(uint64_t)a;
  int32_t hi = (int32_t)(a >> 32);
  int32_t t = lo * qinv;
  int32_t m = __mulhi(t, (int32_t)q);
  return hi - m;
}

__device__ __forceinline__ uint32_t mont_reduce(int64_t a, uint32_t q,
                                                int32_t qinv) {
  int32_t r = mont_reduce_lazy(a, q, qinv);
  r += (int32_t)q & (r >> 31);
  return (uint32_t)r;
}

// a, b in (-2^31, 2^31) with |a*b| < q*2^31: a * b * 2^-32 mod q in (-q,q)
__device__ __forceinline__ int32_t mont_mul_lazy(int32_t a, int32_t b,
                                                 uint32_t q, int32_t qinv) {
  return mont_reduce_lazy((int64_t)a * (int64_t)b, q, qinv);
}

// a, b in [0, q): a * b * 2^-32 mod q in [0, q)
__device__ __forceinline__ uint32_t mont_mul(uint32_t a, uint32_t b,
                                             uint32_t q, int32_t qinv) {
  int32_t r = mont_mul_lazy((int32_t)a, (int32_t)b, q, qinv);
  r += (int32_t)q & (r >> 31);
  return (uint32_t)r;
}

// Montgomery product with a constant w whose companion wq = w * qinv mod
// 2^32 is precomputed: three multiplies instead of four.
__device__ __forceinline__ int32_t mont_mul_pre(int32_t a, int32_t w,
                                                int32_t wq, uint32_t q) {
  int32_t hi = __mulhi(a, w);
  int32_t t = a * wq;
  return hi - __mulhi(t, (int32_t)q);
}

__device__ __forceinline__ uint32_t add_mod(uint32_t a, uint32_t b,
                                            uint32_t q) {
  uint32_t r = a + b;
  return r >= q ? r - q : r;
}

__device__ __forceinline__ uint32_t sub_mod(uint32_t a, uint32_t b,
                                            uint32_t q) {
  int32_t r = (int32_t)a - (int32_t)b;
  r += (int32_t)q & (r >> 31);
  return (uint32_t)r;
}

// [0, q) -> [-(q-1)/2, (q-1)/2]  (Basic.cuh:226-239)
__device__ __forceinline__ int32_t centre(uint32_t a, uint32_t q) {
  return a > (q >> 1) ? (int32_t)a - (int32_t)q : (int32_t)a;
}

// ---------------------------------------------------------------------------
// Exact FP64 modular arithmetic
// ---------------------------------------------------------------------------

struct ModF {
  double q;     // the prime
  double qinv;  // 1/q rounded
};

// a * w mod q for integer-valued doubles; |a| < 2^50, 0 <= w < q < 2^31,
// wq = fl(w qr) with qr = 1/q rounded. Result r == a*w (mod q) exactly with
// |r| < q/2 + 3 |a w| 2^-53 (< q/2 + 2^20 for |a| < 2^40): k (three
// roundings) is within 1/2 + 3 |a w / q| 2^-53 of a w / q, hi - k q is an
// integer small enough to be exact in the fma, and lo is the exact
// rounding error of hi. tests/cpp/test_device_arith.cpp checks this and
// every other function here against 128-bit integer arithmetic.
__device__ __forceinline__ double mulmod_f(double a, double w, double wq,
                                           double q) {
  double hi = a * w;
  double lo = __builtin_fma(a, w, -hi);
  double k = __builtin_rint(a * wq);
  return __builtin_fma(-k, q, hi) + lo;
}

// mulmod_f for a multiplier that changes from product to product (twiddles,
// key words, data): the quotient comes from the rounded product,
// k = rint(fl(a w) qr), qr = 1/q rounded. That is three roundings, as in
// a * fl(w qr), so k is within 1/2 + 3 |a w / q| 2^-53 of a w / q and
// |r| < q/2 + 3 |a w| 2^-53 (< q/2 + 2^20 for |a| < 2^40); the
// exactness argument of mulmod_f is unchanged, and no w/q is formed.
__device__ __forceinline__ double mulmod_fq(double a, double w, double q,
                                            double qr) {
  double hi = a * w;
  double lo = __builtin_fma(a, w, -hi);
  double k = __builtin_rint(hi * qr);
  return __builtin_fma(-k, q, hi) + lo;
}

// |x| < 2^50 -> x mod q in (-0.75 q, 0.75 q) (one rounding step; within
// (-q/2 - 1, q/2 + 1) for |x| < 2^40)
__device__ __forceinline__ double reduce_f(double x, const ModF &m) {
  double k = __builtin_rint(x * m.qinv);
  return __builtin_fma(-k, m.q, x);
}

// |r| < q (a mulmod_f or reduce_f result) -> canonical [0, q) as uint32.
// r converts to int32 exactly; as unsigned u, u + q wraps to r + q in
// (0, q) when r < 0 and exceeds u = r otherwise (q < 2^31), so the unsigned
// minimum is the canonical residue.
__device__ __forceinline__ uint32_t canon_small(double r, double q) {
  const uint32_t u = (uint32_t)(int32_t)r;
  const uint32_t v = u + (uint32_t)q;
  return u < v ? u : v;
}

// |x| < 2^50 -> canonical [0, q) as uint32
__device__ __forceinline__ uint32_t canon_f(double x, const ModF &m) {
  return canon_small(reduce_f(x, m), m.q);
}

// |x| < 2^50 -> centred [-(q-1)/2, (q-1)/2] as int32
__device__ __forceinline__ int32_t centre_f(double x, const ModF &m) {
  double r = reduce_f(x, m);
  double h = (m.q - 1.0) * 0.5;
  r = r > h ? r - m.q : r;
  r = r < -h ? r + m.q : r;
  return (int32_t)r;
}

}  // namespace ck
