// Element-wise kernels over RNS polynomials (reference:
// src/core/ElementWise.cu). All are HBM-bound: each thread moves 4
// consecutive coefficients of one limb with 16-byte loads/stores, the limb
// (and so the prime) is uniform per workgroup, and products are Montgomery
// products of stored values exactly as the reference's MultMontgomery
// (include/common/Basic.cuh:251-260), so outputs are bit-identical.
#include "ck_elem.h"

namespace ck {
namespace {

constexpr int kThreads = 256;
constexpr int kVec = 4;
constexpr int kBlocksPerLimb = 65536 / (kThreads * kVec);

__device__ __forceinline__ int ew_prime(const EwNp &np, int y) {
  return y < np.num_q ? np.q_prime0 + y : np.aux_prime0 + (y - np.num_q);
}
__device__ __forceinline__ size_t ew_off(const EwOperand &o, const EwNp &np,
                                         int y, int k) {
  return (size_t)(y < np.num_q ? y : y + o.gap) * 65536 + k;
}

#define EW_PROLOGUE                                    \
  const int y = blockIdx.y;                            \
  const int k = (blockIdx.x * kThreads + threadIdx.x) * kVec; \
  const int p = ew_prime(np, y);                       \
  const uint32_t q = P.q[p];                           \
  const int32_t qinv = P.qinv[p];                      \
  (void)qinv;

// every operand is device memory: global (not flat) accesses, also through
// pointers that a batched plan's table holds
// (a native vector type: HIP's uint4 copy goes through a generic reference)
typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) v4u32 gv4u32;
__device__ __forceinline__ uint4 ld4(const uint32_t *p) {
  const v4u32 v = *(const gv4u32 *)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st4(uint32_t *p, uint4 v) {
  const v4u32 w = {v.x, v.y, v.z, v.w};
  *(gv4u32 *)p = w;
}
__device__ __forceinline__ uint32_t ld1(const uint32_t *p) {
  return *(const __attribute__((address_space(1))) uint32_t *)p;
}
__device__ __forceinline__ void st1(uint32_t *p, uint32_t v) {
  *(__attribute__((address_space(1))) uint32_t *)p = v;
}

// compile-time count tag for the per-count bodies below
template <int N>
struct IntC {
  static constexpr int value = N;
};

template <typename F>
__device__ __forceinline__ uint4 map2(uint4 a, uint4 b, F f) {
  return make_uint4(f(a.x, b.x), f(a.y, b.y), f(a.z, b.z), f(a.w, b.w));
}

// op: 0 add, 1 sub, 2 mult (Montgomery)
template <int OP>
__global__ __launch_bounds__(kThreads) void binary_kernel(
    const EwNp np, const PrimeTab P, EwOperand x, EwOperand yv, EwOperand out,
    int npoly) {
  EW_PROLOGUE
  for (int j = 0; j < npoly; j++) {
    uint4 a = ld4(x.ptr[j] + ew_off(x, np, y, k));
    uint4 b = ld4(yv.ptr[j] + ew_off(yv, np, y, k));
    uint4 r;
    if (OP == 0)
      r = map2(a, b, [&](uint32_t u, uint32_t v) { return add_mod(u, v, q); });
    else if (OP == 1)
      r = map2(a, b, [&](uint32_t u, uint32_t v) { return sub_mod(u, v, q); });
    else
      r = map2(a, b,
               [&](uint32_t u, uint32_t v) { return mont_mul(u, v, q, qinv); });
    st4(out.ptr[j] + ew_off(out, np, y, k), r);
  }
}

// op: 0 neg, 1 copy
template <int OP>
__global__ __launch_bounds__(kThreads) void unary_kernel(const EwNp np,
                                                         const PrimeTab P,
                                                         EwOperand x,
                                                         EwOperand out,
                                                         int npoly) {
  EW_PROLOGUE
  for (int j = 0; j < npoly; j++) {
    uint4 a = ld4(x.ptr[j] + ew_off(x, np, y, k));
    uint4 r = a;
    if (OP == 0)
      r = map2(a, a, [&](uint32_t u, uint32_t) { return u ? q - u : 0u; });
    st4(out.ptr[j] + ew_off(out, np, y, k), r);
  }
}

// CPAccum<const> with one term = MultConst: x * c_y (c in Montgomery form)
// op: 0 mult, 1 add (c + x), 2 sub (x - c), 3 opposite sub (c - x)
// (AddConst / SubConst / SubOppositeConst, ElementWise.cu:145-223)
template <int OP>
__global__ __launch_bounds__(kThreads) void const_kernel(
    const EwNp np, const PrimeTab P, EwOperand x, const uint32_t *consts,
    int const_gap, EwOperand out, int npoly) {
  EW_PROLOGUE
  const uint32_t c = consts[y < np.num_q ? y : y + const_gap];
  for (int j = 0; j < npoly; j++) {
    uint4 a = ld4(x.ptr[j] + ew_off(x, np, y, k));
    uint4 r;
    if (OP == 0)
      r = map2(a, a, [&](uint32_t u, uint32_t) { return mont_mul(u, c, q, qinv); });
    else if (OP == 1)
      r = map2(a, a, [&](uint32_t u, uint32_t) { return add_mod(u, c, q); });
    else if (OP == 2)
      r = map2(a, a, [&](uint32_t u, uint32_t) { return sub_mod(u, c, q); });
    else
      r = map2(a, a, [&](uint32_t u, uint32_t) { return sub_mod(c, u, q); });
    st4(out.ptr[j] + ew_off(out, np, y, k), r);
  }
}

// MultImaginaryUnit (ElementWise.cu:625-658): multiplication by X^(N/2).
// NTT position k holds the value at psi^(2 brv(k) + 1), so the factor is
// psi^(N/2) (-1)^brv(k), and brv(k) is odd exactly for k >= N/2. imag[p]
// is psi_p^(N/2) in Montgomery form.
__global__ __launch_bounds__(kThreads) void imag_unit_kernel(
    const EwNp np, const PrimeTab P, const uint32_t *imag, EwOperand x,
    EwOperand out, int npoly) {
  EW_PROLOGUE
  const uint32_t c0 = imag[p];
  const uint32_t c = k >= 32768 ? q - c0 : c0;
  for (int j = 0; j < npoly; j++) {
    uint4 a = ld4(x.ptr[j] + ew_off(x, np, y, k));
    uint4 r = map2(a, a, [&](uint32_t u, uint32_t) { return mont_mul(u, c, q, qinv); });
    st4(out.ptr[j] + ew_off(out, np, y, k), r);
  }
}

// Sum of Montgomery products (see AccumArgs). Each product is canonical
// and the sum is reduced after every term, so the result is the canonical
// residue whatever the order, as the reference's kernels produce it.
__global__ __launch_bounds__(kThreads) void accum_kernel(const EwNp np,
                                                         const PrimeTab P,
                                                         const AccumArgs A) {
  EW_PROLOGUE
  const size_t off = (size_t)y * 65536 + k;
  for (int j = 0; j < A.npoly; j++) {
    uint4 acc = A.accumulate ? ld4(A.dst[j] + off) : make_uint4(0, 0, 0, 0);
    for (int t = 0; t < A.nterms; t++) {
      const uint4 a = ld4(A.x[t][j] + off);
      uint4 prod;
      if (A.y_const) {
        const uint32_t c = A.y[t][y];
        prod = map2(a, a, [&](uint32_t u, uint32_t) { return mont_mul(u, c, q, qinv); });
      } else {
        prod = map2(a, ld4(A.y[t] + off),
                    [&](uint32_t u, uint32_t v) { return mont_mul(u, v, q, qinv); });
      }
      acc = map2(acc, prod, [&](uint32_t u, uint32_t v) { return add_mod(u, v, q); });
    }
    st4(A.dst[j] + off, acc);
  }
}

// Giant-step accumulation (GsAccumArgs). NB (>= nbs) bounds the baby steps
// held in registers, VEC the coefficients per thread. Every load is
// unconditional (a missing plaintext or a baby step past nbs reads a valid
// operand and is masked to zero afterwards), so the loads of a giant step
// are all in flight before its first product.
//
// Arithmetic in exact FP64 instead of 32-bit Montgomery products (whose
// multiplies issue at a quarter of the FP64 rate): with the plaintext word
// split as m = mh 2^14 + ml, the sums Sh = sum x mh (< 16 2^31 2^17 = 2^52)
// and Sl = sum x ml (< 2^49) over at most 16 terms are exact doubles, and
//   sum_t x_t (*) m_t = R^-1 (Sh 2^14 + Sl)  mod q
// is one reduction of Sh 2^14 (a power-of-two scaling, exact), one FP64
// modular product by R^-1 (its operand below 2^49 + q) and a canonical
// fix-up per output.
template <int NB, int VEC>
__global__ __launch_bounds__(kThreads) void gs_accum_kernel(const EwNp np,
                                                            const PrimeTab P,
                                                            const GsAccumArgs A) {
  typedef uint32_t vec_t __attribute__((ext_vector_type(VEC)));
  typedef __attribute__((address_space(1))) vec_t gvec_t;
  const int y = blockIdx.y;
  const int k = (blockIdx.x * kThreads + threadIdx.x) * VEC;
  const int p = ew_prime(np, y);
  const double q = P.qd[p], qr = P.qr[p];
  const double rinv = (double)mont_reduce(1, P.q[p], P.qinv[p]), rinvq = rinv * qr;
  const size_t off = (size_t)y * 65536 + k;
  vec_t xb[NB], xa[NB];
#pragma unroll
  for (int t = 0; t < NB; t++) {
    const int tt = t < A.nbs ? t : 0;
    xb[t] = *(const gvec_t *)(A.x[tt][0] + off);
    xa[t] = *(const gvec_t *)(A.x[tt][1] + off);
  }
  for (int g = 0; g < A.ngs; g++) {
    vec_t m[NB];
#pragma unroll
    for (int t = 0; t < NB; t++) {
      const uint32_t *pp = A.pt[g][t < A.nbs ? t : 0];
      const bool use = t < A.nbs && pp != nullptr;
      const vec_t v = *(const gvec_t *)((use ? pp : A.x[0][0]) + off);
      m[t] = use ? v : (vec_t)0;
    }
    double hb[VEC], lb[VEC], ha[VEC], la[VEC];
#pragma unroll
    for (int e = 0; e < VEC; e++) hb[e] = lb[e] = ha[e] = la[e] = 0.0;
#pragma unroll
    for (int t = 0; t < NB; t++)
#pragma unroll
      for (int e = 0; e < VEC; e++) {
        const double mh = (double)(m[t][e] >> 14), ml = (double)(m[t][e] & 0x3FFFu);
        const double vb = (double)xb[t][e], va = (double)xa[t][e];
        hb[e] = __builtin_fma(vb, mh, hb[e]);
        lb[e] = __builtin_fma(vb, ml, lb[e]);
        ha[e] = __builtin_fma(va, mh, ha[e]);
        la[e] = __builtin_fma(va, ml, la[e]);
      }
    vec_t rb, ra;
#pragma unroll
    for (int e = 0; e < VEC; e++) {
      const double bigb = hb[e] * 16384.0, biga = ha[e] * 16384.0;  // exact
      const double kb = __builtin_rint(bigb * qr), ka = __builtin_rint(biga * qr);
      const double sb = __builtin_fma(-kb, q, bigb) + lb[e];
      const double sa = __builtin_fma(-ka, q, biga) + la[e];
      rb[e] = canon_small(mulmod_f(sb, rinv, rinvq, q), q);
      ra[e] = canon_small(mulmod_f(sa, rinv, rinvq, q), q);
    }
    *(gvec_t *)(A.dst[g][0] + off) = rb;
    *(gvec_t *)(A.dst[g][1] + off) = ra;
  }
}

// ModUpToMax1 / ModUpToMax2 (ElementWise.cu:579-623), generalised to any
// source and destination primes: coefficient-domain limbs of a one- or
// two-prime base lifted to every destination prime. Two primes: v = s0 q1
// + s1 q0 in [0, 2 q0 q1), centred into (-q0 q1 / 2, q0 q1 / 2], then
// reduced; one prime: s0 mod p (not centred). With s_i = [x (q0 q1 /
// q_i)^-1]_{q_i} this is the exact CRT lift of x.
__global__ __launch_bounds__(kThreads) void crt_lift_kernel(const PrimeTab P,
                                                            const uint32_t *src,
                                                            const LiftArgs A,
                                                            uint32_t *dst) {
  const int y = blockIdx.y;
  const int k = (blockIdx.x * kThreads + threadIdx.x) * kVec;
  const uint32_t q = P.q[A.dst_prime[y]];
  const uint4 s0 = ld4(src + k);
  uint4 r;
  if (A.nsrc == 2) {
    const uint4 s1 = ld4(src + 65536 + k);
    const int64_t qq = (int64_t)A.q0 * A.q1, half = qq >> 1;
    auto lift = [&](uint32_t u, uint32_t v) {
      int64_t t = (int64_t)u * A.q1 + (int64_t)v * A.q0;
      if (t >= qq) t -= qq;
      if (t > half) t -= qq;
      int64_t m = t % (int64_t)q;
      if (m < 0) m += q;
      return (uint32_t)m;
    };
    r = make_uint4(lift(s0.x, s1.x), lift(s0.y, s1.y), lift(s0.z, s1.z),
                   lift(s0.w, s1.w));
  } else {
    r = make_uint4(s0.x % q, s0.y % q, s0.z % q, s0.w % q);
  }
  st4(dst + (size_t)y * 65536 + k, r);
}

// Tensor (ElementWise.cu:504-546): (b1 b2, (b1+a1)(b2+a2) - b1 b2 - a1 a2,
// a1 a2) with the Karatsuba middle term on signed lazy sums.
__global__ __launch_bounds__(kThreads) void tensor_kernel(
    const EwNp np, const PrimeTab P, EwOperand c1, EwOperand c2,
    EwOperand out) {
  EW_PROLOGUE
  uint4 b1 = ld4(c1.ptr[0] + ew_off(c1, np, y, k));
  uint4 a1 = ld4(c1.ptr[1] + ew_off(c1, np, y, k));
  uint4 b2 = ld4(c2.ptr[0] + ew_off(c2, np, y, k));
  uint4 a2 = ld4(c2.ptr[1] + ew_off(c2, np, y, k));
  uint32_t bb[4], aa[4], mid[4];
  const uint32_t B1[4] = {b1.x, b1.y, b1.z, b1.w};
  const uint32_t A1[4] = {a1.x, a1.y, a1.z, a1.w};
  const uint32_t B2[4] = {b2.x, b2.y, b2.z, b2.w};
  const uint32_t A2[4] = {a2.x, a2.y, a2.z, a2.w};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    bb[i] = mont_mul(B1[i], B2[i], q, qinv);
    aa[i] = mont_mul(A1[i], A2[i], q, qinv);
    const int32_t s1 = (int32_t)(B1[i] - q) + (int32_t)A1[i];  // [-q, q-2]
    const int32_t s2 = (int32_t)(B2[i] - q) + (int32_t)A2[i];
    uint32_t m = mont_reduce((int64_t)s1 * (int64_t)s2, q, qinv);
    m = sub_mod(m, bb[i], q);
    mid[i] = sub_mod(m, aa[i], q);
  }
  const size_t o = ew_off(out, np, y, k);
  st4(out.ptr[0] + o, make_uint4(bb[0], bb[1], bb[2], bb[3]));
  st4(out.ptr[1] + o, make_uint4(mid[0], mid[1], mid[2], mid[3]));
  st4(out.ptr[2] + o, make_uint4(aa[0], aa[1], aa[2], aa[3]));
}

// TensorSquare (ElementWise.cu:549-575): the Tensor of a ciphertext with
// itself reads its two polynomials once: (b^2, 2 b a, a^2), Montgomery
// products as in tensor_kernel (2 b a equals the Karatsuba middle term
// (b + a)^2 - b^2 - a^2 as a canonical residue).
__global__ __launch_bounds__(kThreads) void tensor_square_kernel(
    const EwNp np, const PrimeTab P, EwOperand c1, EwOperand out) {
  EW_PROLOGUE
  const uint4 b = ld4(c1.ptr[0] + ew_off(c1, np, y, k));
  const uint4 a = ld4(c1.ptr[1] + ew_off(c1, np, y, k));
  const uint4 bb = map2(b, b, [&](uint32_t u, uint32_t v) { return mont_mul(u, v, q, qinv); });
  const uint4 ba = map2(b, a, [&](uint32_t u, uint32_t v) { return mont_mul(u, v, q, qinv); });
  const uint4 aa = map2(a, a, [&](uint32_t u, uint32_t v) { return mont_mul(u, v, q, qinv); });
  const uint4 mid = map2(ba, ba, [&](uint32_t u, uint32_t v) { return add_mod(u, v, q); });
  const size_t o = ew_off(out, np, y, k);
  st4(out.ptr[0] + o, bb);
  st4(out.ptr[1] + o, mid);
  st4(out.ptr[2] + o, aa);
}

// Key inner product of MultKeyNoModDown (Context.cpp:609-712 ->
// ElementWise::PAccum, CPAccum<2,false>) fused with the P * (b, a) fold
// (CAccum, CPAccumAdd<2,true>):
//   acc_x[y] = sum_g modup_g[y] (*) key_x_g[prime(y)]  (+ P (*) add_x[y])
__device__ __forceinline__ void key_accum_body(const EwNp &np,
                                               const PrimeTab &P,
                                               const KeyAccumArgs &A, int y,
                                               int k) {
  const int p = ew_prime(np, y);
  const uint32_t q = P.q[p];
  const int32_t qinv = P.qinv[p];
  const size_t tot_stride = (size_t)np.num_total * 65536;
  uint32_t rb[4] = {0, 0, 0, 0}, ra[4] = {0, 0, 0, 0};
  const size_t my = (size_t)y * 65536 + k;
  const int kl = y < np.num_q ? A.key_q0 + y : A.key_aux0 + (y - np.num_q);
  const size_t kp = (size_t)kl * 65536 + k;
  // HMult form: Tensor (ElementWise.cu:504-546) of the two input ciphertexts
  // on this q limb
  uint32_t d0[4] = {0, 0, 0, 0}, d1[4] = {0, 0, 0, 0}, d2[4] = {0, 0, 0, 0};
  const bool tens = A.use_tensor && y < np.num_q;
  if (tens) {
    const uint4 b1 = ld4(A.tensor[0] + my), a1 = ld4(A.tensor[1] + my);
    const uint4 b2 = ld4(A.tensor[2] + my), a2 = ld4(A.tensor[3] + my);
    const uint32_t B1[4] = {b1.x, b1.y, b1.z, b1.w};
    const uint32_t A1[4] = {a1.x, a1.y, a1.z, a1.w};
    const uint32_t B2[4] = {b2.x, b2.y, b2.z, b2.w};
    const uint32_t A2[4] = {a2.x, a2.y, a2.z, a2.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      d0[i] = mont_mul(B1[i], B2[i], q, qinv);
      d1[i] = add_mod(mont_mul(B1[i], A2[i], q, qinv),
                      mont_mul(A1[i], B2[i], q, qinv), q);
      d2[i] = mont_mul(A1[i], A2[i], q, qinv);
    }
  }
  // one group at a time: the kernel is HBM-bound with many waves per SIMD,
  // which hide each group's load latency; batching the groups' loads (more
  // registers, fewer waves) measured slower
  for (int gi = 0; gi < A.ngroups; gi++) {
    const int g = A.group[gi];
    const bool own = (y >= A.own0[g] && y < A.own1[g]);
    const uint32_t *src = own ? (tens ? A.tensor[1] : A.x)
                              : A.modup + (size_t)g * tot_stride;
    uint4 mu = ld4(src + my);
    const uint4 kb = ld4(A.key_b[g] + kp);
    const uint4 ka = ld4(A.key_a[g] + kp);
    if (own && tens) mu = make_uint4(d2[0], d2[1], d2[2], d2[3]);
    const uint32_t M[4] = {mu.x, mu.y, mu.z, mu.w};
    const uint32_t KB[4] = {kb.x, kb.y, kb.z, kb.w};
    const uint32_t KA[4] = {ka.x, ka.y, ka.z, ka.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      rb[i] = add_mod(rb[i], mont_mul(M[i], KB[i], q, qinv), q);
      ra[i] = add_mod(ra[i], mont_mul(M[i], KA[i], q, qinv), q);
    }
  }
  if (tens) {
    const uint32_t pp = ld1(A.p_prod + p);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      rb[i] = add_mod(rb[i], mont_mul(pp, d0[i], q, qinv), q);
      ra[i] = add_mod(ra[i], mont_mul(pp, d1[i], q, qinv), q);
    }
  } else if (y < np.num_q) {
    const uint32_t pp = ld1(A.p_prod + p);
    if (A.add_b) {
      uint4 v = ld4(A.add_b + my);
      const uint32_t V[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int i = 0; i < 4; i++)
        rb[i] = add_mod(rb[i], mont_mul(pp, V[i], q, qinv), q);
    }
    if (A.add_a) {
      uint4 v = ld4(A.add_a + my);
      const uint32_t V[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int i = 0; i < 4; i++)
        ra[i] = add_mod(ra[i], mont_mul(pp, V[i], q, qinv), q);
    }
  }
  st4(A.out_b + my, make_uint4(rb[0], rb[1], rb[2], rb[3]));
  st4(A.out_a + my, make_uint4(ra[0], ra[1], ra[2], ra[3]));
}

__global__ __launch_bounds__(kThreads) void key_accum_kernel(
    const EwNp np, const PrimeTab P, const KeyAccumArgs A) {
  key_accum_body(np, P, A, blockIdx.y, (blockIdx.x * kThreads + threadIdx.x) * kVec);
}

__global__ __launch_bounds__(kThreads) void key_accum_table_kernel(
    const KeyAccumItem *__restrict__ items, const EwJob *__restrict__ jobs,
    const PrimeTab P) {
  const EwJob jb = jobs[blockIdx.y];
  const KeyAccumItem &it = items[jb.item];
  key_accum_body(it.np, P, it.A, jb.y,
                 (blockIdx.x * kThreads + threadIdx.x) * kVec);
}

// Hoisted baby step (BSFusedKernel, src/extension/Hoist.cu:21-78): the ModUp
// values of 4 coefficients are loaded once and reused by every rotation;
// each rotation's products are accumulated and scattered to the
// automorphism's positions d = brv17((brv17(x) + 1) g^-1 - 1). The grid is
// XCD-aware: the 64 workgroups of one limb share an XCD (workgroup i runs on
// XCD i mod 8), so the partial-line scatter writes of a limb merge in one
// L2.
constexpr int kBsBlocksPerLimb = 65536 / (kThreads * kVec);

__device__ __forceinline__ void bs_fused_body(const EwNp &np, const PrimeTab &P,
                                              const BsFusedArgs &A, int y,
                                              int k) {
  const int p = ew_prime(np, y);
  const uint32_t q = P.q[p];
  const int32_t qinv = P.qinv[p];
  const size_t my = (size_t)y * 65536 + k;
  const size_t kp = (size_t)p * 65536 + k;
  const size_t tot_stride = (size_t)np.num_total * 65536;
  // register arrays: every loop over groups is unrolled to the static bound
  // loads of all groups issued before any use (a load in a branch that
  // also uses it is waited for at once): exact bodies per count up to 4
  uint32_t mu[kMaxBsGroups][4];
  auto load_mu = [&](auto ng_tag, int g0) {
    constexpr int NG = decltype(ng_tag)::value;
    uint4 v[NG];
#pragma unroll
    for (int b = 0; b < NG; b++) {
      const int gi = g0 + b;
      const bool own = (y >= A.own0[gi] && y < A.own1[gi]);
      const uint32_t *src = own ? A.x : A.modup + (size_t)A.group[gi] * tot_stride;
      v[b] = ld4(src + my);
    }
#pragma unroll
    for (int b = 0; b < NG; b++) {
      mu[g0 + b][0] = v[b].x;
      mu[g0 + b][1] = v[b].y;
      mu[g0 + b][2] = v[b].z;
      mu[g0 + b][3] = v[b].w;
    }
  };
  const int ng = A.ngroups;
  const int nfull = ng >= 4 ? 4 : 0;  // kMaxBsGroups = 8: at most two batches
  if (nfull) load_mu(IntC<4>{}, 0);
  switch (ng - nfull) {
    case 1: load_mu(IntC<1>{}, nfull); break;
    case 2: load_mu(IntC<2>{}, nfull); break;
    case 3: load_mu(IntC<3>{}, nfull); break;
    case 4: load_mu(IntC<4>{}, nfull); break;
    default: break;
  }
  uint32_t pb[4] = {0, 0, 0, 0};
  if (y < np.num_q) {
    const uint32_t pp = ld1(A.p_prod + p);
    const uint4 v = ld4(A.b + my);
    pb[0] = mont_mul(pp, v.x, q, qinv);
    pb[1] = mont_mul(pp, v.y, q, qinv);
    pb[2] = mont_mul(pp, v.z, q, qinv);
    pb[3] = mont_mul(pp, v.w, q, qinv);
  }
  uint32_t d[4];
  for (int r = 0; r < A.nrot; r++) {
    uint32_t rb[4] = {pb[0], pb[1], pb[2], pb[3]}, ra[4] = {0, 0, 0, 0};
    auto keys = [&](auto ng_tag, int g0) {
      constexpr int NG = decltype(ng_tag)::value;
      uint4 kb[NG], ka[NG];
#pragma unroll
      for (int b = 0; b < NG; b++) {
        kb[b] = ld4(A.key_b[r][g0 + b] + kp);
        ka[b] = ld4(A.key_a[r][g0 + b] + kp);
      }
#pragma unroll
      for (int b = 0; b < NG; b++) {
        const uint32_t KB[4] = {kb[b].x, kb[b].y, kb[b].z, kb[b].w};
        const uint32_t KA[4] = {ka[b].x, ka[b].y, ka[b].z, ka[b].w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
          rb[i] = add_mod(rb[i], mont_mul(mu[g0 + b][i], KB[i], q, qinv), q);
          ra[i] = add_mod(ra[i], mont_mul(mu[g0 + b][i], KA[i], q, qinv), q);
        }
      }
    };
    if (nfull) keys(IntC<4>{}, 0);
    switch (ng - nfull) {
      case 1: keys(IntC<1>{}, nfull); break;
      case 2: keys(IntC<2>{}, nfull); break;
      case 3: keys(IntC<3>{}, nfull); break;
      case 4: keys(IntC<4>{}, nfull); break;
      default: break;
    }
    const uint32_t g = A.ginv[r];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const uint32_t e = (__brev((uint32_t)(k + i)) >> 15) + 1u;
      d[i] = __brev((e * g - 1u) & 0x1FFFFu) >> 15;
    }
    uint32_t *ob = A.out_b[r] + (size_t)y * 65536;
    uint32_t *oa = A.out_a[r] + (size_t)y * 65536;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      st1(ob + d[i], rb[i]);
      st1(oa + d[i], ra[i]);
    }
  }
}

// limb slot of workgroup blockIdx.x for `nlimbs` limbs (XCD-aware order);
// returns -1 for the padding workgroups of the last group of 8 limbs
__device__ __forceinline__ int bs_limb(int nlimbs, int &chunk) {
  const int xcd = blockIdx.x & 7;
  const int j = blockIdx.x >> 3;
  chunk = j % kBsBlocksPerLimb;
  const int y = (j / kBsBlocksPerLimb) * 8 + xcd;
  return y < nlimbs ? y : -1;
}

__global__ __launch_bounds__(kThreads) void bs_fused_kernel(
    const EwNp np, const PrimeTab P, const BsFusedArgs A) {
  int chunk;
  const int y = bs_limb(np.num_total, chunk);
  if (y < 0) return;
  bs_fused_body(np, P, A, y, (chunk * kThreads + threadIdx.x) * kVec);
}

__global__ __launch_bounds__(kThreads) void bs_fused_table_kernel(
    const BsFusedItem *__restrict__ items, const EwJob *__restrict__ jobs,
    int njobs, const PrimeTab P) {
  int chunk;
  const int j = bs_limb(njobs, chunk);
  if (j < 0) return;
  const EwJob jb = jobs[j];
  const BsFusedItem &it = items[jb.item];
  bs_fused_body(it.np, P, it.A, jb.y, (chunk * kThreads + threadIdx.x) * kVec);
}

// Permute (ElementWise.cu:370-391): dst[i] = src[brv17((brv17(i)+1) g - 1)]
// (log N + 1 = 17 bits; wraps mod 2N). Gathers are L2-served; writes are
// coalesced.
__global__ __launch_bounds__(kThreads) void permute_kernel(
    const EwNp np, EwOperand x, EwOperand out, uint32_t g, int npoly) {
  const int y = blockIdx.y;
  const int k = blockIdx.x * kThreads + threadIdx.x;
  const uint32_t e = (__brev((uint32_t)k) >> 15) + 1u;
  const uint32_t s = __brev((e * g - 1u) & 0x1FFFFu) >> 15;
  for (int j = 0; j < npoly; j++) {
    const uint32_t *src = x.ptr[j] + (size_t)(y < np.num_q ? y : y + x.gap) * 65536;
    out.ptr[j][(size_t)(y < np.num_q ? y : y + out.gap) * 65536 + k] = src[s];
  }
}

// PermAccArgs: the NS (>= nsrc) source values of each polynomial are
// loaded (a null source reads dst[0] and is masked to zero) before the
// first addition; one body per source bound keeps dummy loads few. A
// masked term past nsrc gathers with the identity index (g = 1 gives
// idx = k), so its dummy load stays inside the limb also when nsrc = 0
// (an unset g = 0 would give e g - 1 = -1, index 2^17 - 1).
template <int NS>
__global__ __launch_bounds__(kThreads) void permute_accum_kernel(const EwNp np,
                                                                 const PrimeTab P,
                                                                 const PermAccArgs A) {
  const int y = blockIdx.y;
  const int k = blockIdx.x * kThreads + threadIdx.x;
  const uint32_t q = P.q[ew_prime(np, y)];
  const size_t base = (size_t)y * 65536;
  const uint32_t e = (__brev((uint32_t)k) >> 15) + 1u;
  uint32_t v[NS][2];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int ss = s < A.nsrc ? s : 0;
    const uint32_t g = s < A.nsrc ? A.g[ss] : 1u;
    const uint32_t idx = g == 1u ? (uint32_t)k : __brev((e * g - 1u) & 0x1FFFFu) >> 15;
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const uint32_t *p = A.src[ss][j < A.npoly ? j : 0];
      const bool use = s < A.nsrc && j < A.npoly && p != nullptr;
      const uint32_t w = ld1((use ? p : A.dst[0]) + base + idx);
      v[s][j] = use ? w : 0u;
    }
  }
#pragma unroll
  for (int j = 0; j < 2; j++) {
    if (j >= A.npoly) break;
    uint32_t acc = A.accumulate ? ld1(A.dst[j] + base + k) : 0u;
#pragma unroll
    for (int s = 0; s < NS; s++) acc = add_mod(acc, v[s][j], q);
    st1(A.dst[j] + base + k, acc);
  }
}

}  // namespace

static dim3 ew_grid(const EwNp &np) { return dim3(kBlocksPerLimb, np.num_total); }

void launch_binary(const Context &ctx, int op, const EwNp &np,
                   const EwOperand &x, const EwOperand &y, const EwOperand &out,
                   int npoly, hipStream_t st) {
  if (np.num_total <= 0) return;
  const int limbs = np.num_total * npoly;
  if (op == 0)
    launch(ctx, "add", limbs, binary_kernel<0>, ew_grid(np), dim3(kThreads), st, np,
           ctx.ptab, x, y, out, npoly);
  else if (op == 1)
    launch(ctx, "sub", limbs, binary_kernel<1>, ew_grid(np), dim3(kThreads), st, np,
           ctx.ptab, x, y, out, npoly);
  else
    launch(ctx, "mult", limbs, binary_kernel<2>, ew_grid(np), dim3(kThreads), st, np,
           ctx.ptab, x, y, out, npoly);
}

void launch_unary(const Context &ctx, int op, const EwNp &np,
                  const EwOperand &x, const EwOperand &out, int npoly,
                  hipStream_t st) {
  if (np.num_total <= 0) return;
  const int limbs = np.num_total * npoly;
  if (op == 0)
    launch(ctx, "neg", limbs, unary_kernel<0>, ew_grid(np), dim3(kThreads), st, np,
           ctx.ptab, x, out, npoly);
  else
    launch(ctx, "copy", limbs, unary_kernel<1>, ew_grid(np), dim3(kThreads), st, np,
           ctx.ptab, x, out, npoly);
}

void launch_const(const Context &ctx, int op, const EwNp &np,
                  const EwOperand &x, const uint32_t *consts, int const_gap,
                  const EwOperand &out, int npoly, hipStream_t st) {
  if (np.num_total <= 0) return;
  const int limbs = np.num_total * npoly;
  if (op == 0)
    launch(ctx, "mult_const", limbs, const_kernel<0>, ew_grid(np), dim3(kThreads), st,
           np, ctx.ptab, x, consts, const_gap, out, npoly);
  else if (op == 1)
    launch(ctx, "add_const", limbs, const_kernel<1>, ew_grid(np), dim3(kThreads), st,
           np, ctx.ptab, x, consts, const_gap, out, npoly);
  else if (op == 2)
    launch(ctx, "sub_const", limbs, const_kernel<2>, ew_grid(np), dim3(kThreads), st,
           np, ctx.ptab, x, consts, const_gap, out, npoly);
  else
    launch(ctx, "const_sub", limbs, const_kernel<3>, ew_grid(np), dim3(kThreads), st,
           np, ctx.ptab, x, consts, const_gap, out, npoly);
}

void launch_imag_unit(const Context &ctx, const EwNp &np, const EwOperand &x,
                      const EwOperand &out, int npoly, hipStream_t st) {
  if (np.num_total <= 0) return;
  launch(ctx, "mult_imaginary_unit", np.num_total, imag_unit_kernel, ew_grid(np),
         dim3(kThreads), st, np, ctx.ptab, (const uint32_t *)ctx.d_imag, x, out,
         npoly);
}

void launch_accum(const Context &ctx, const EwNp &np, const AccumArgs &A,
                  hipStream_t st) {
  if (np.num_total <= 0) return;
  launch(ctx, A.y_const ? "caccum" : "paccum", np.num_total, accum_kernel, ew_grid(np),
         dim3(kThreads), st, np, ctx.ptab, A);
}

void launch_crt_lift(const Context &ctx, const uint32_t *src, const LiftArgs &A,
                     int ndst, uint32_t *dst, hipStream_t st) {
  if (ndst <= 0) return;
  launch(ctx, "crt_lift", ndst, crt_lift_kernel, dim3(kBlocksPerLimb, ndst),
         dim3(kThreads), st, ctx.ptab, src, A, dst);
}

void launch_tensor(const Context &ctx, const EwNp &np, const EwOperand &c1,
                   const EwOperand &c2, const EwOperand &out, hipStream_t st) {
  if (np.num_total <= 0) return;
  // the same ciphertext twice: the squaring form (ElementWise.cu:940-945)
  if (c1.ptr[0] == c2.ptr[0] && c1.ptr[1] == c2.ptr[1] && c1.gap == c2.gap)
    launch(ctx, "tensor_square", np.num_total, tensor_square_kernel, ew_grid(np),
           dim3(kThreads), st, np, ctx.ptab, c1, out);
  else
    launch(ctx, "tensor", np.num_total, tensor_kernel, ew_grid(np),
           dim3(kThreads), st, np, ctx.ptab, c1, c2, out);
}

void launch_gs_accum(const Context &ctx, const EwNp &np, const GsAccumArgs &A,
                     hipStream_t st) {
  if (A.ngs <= 0 || np.num_total <= 0) return;
  const dim3 grid(65536 / (kThreads * 2), np.num_total);
  if (A.nbs <= 4)
    launch(ctx, "gs_accum", np.num_total, gs_accum_kernel<4, 2>, grid, dim3(kThreads), st,
           np, ctx.ptab, A);
  else if (A.nbs <= 8)
    launch(ctx, "gs_accum", np.num_total, gs_accum_kernel<8, 2>, grid, dim3(kThreads), st,
           np, ctx.ptab, A);
  else
    launch(ctx, "gs_accum", np.num_total, gs_accum_kernel<kMaxGsBaby, 2>, grid,
           dim3(kThreads), st, np, ctx.ptab, A);
}

void launch_key_accum(const Context &ctx, const EwNp &np,
                      const KeyAccumArgs &A, hipStream_t st) {
  if (np.num_total <= 0) return;
  launch(ctx, "key_accum", np.num_total, key_accum_kernel, ew_grid(np),
         dim3(kThreads), st, np, ctx.ptab, A);
}

void launch_permute(const Context &ctx, const EwNp &np, const EwOperand &x,
                    const EwOperand &out, uint32_t g, int npoly,
                    hipStream_t st) {
  if (np.num_total <= 0) return;
  launch(ctx, "permute", np.num_total, permute_kernel,
         dim3(65536 / kThreads, np.num_total), dim3(kThreads), st, np, x, out,
         g, npoly);
}

void launch_permute_accum(const Context &ctx, const EwNp &np, const PermAccArgs &A,
                          hipStream_t st) {
  if (np.num_total <= 0) return;
  const dim3 grid(65536 / kThreads, np.num_total);
  const int limbs = np.num_total * A.npoly;
  if (A.nsrc <= 1)
    launch(ctx, "permute_accum", limbs, permute_accum_kernel<1>, grid, dim3(kThreads), st,
           np, ctx.ptab, A);
  else if (A.nsrc == 2)
    launch(ctx, "permute_accum", limbs, permute_accum_kernel<2>, grid, dim3(kThreads), st,
           np, ctx.ptab, A);
  else if (A.nsrc <= 4)
    launch(ctx, "permute_accum", limbs, permute_accum_kernel<4>, grid, dim3(kThreads), st,
           np, ctx.ptab, A);
  else
    launch(ctx, "permute_accum", limbs, permute_accum_kernel<kMaxPermAcc>, grid,
           dim3(kThreads), st, np, ctx.ptab, A);
}

// Device-to-device copy (DeviceVector copies, src/core/DeviceVector.cpp:
// 105-124, and ModUp's copy of a group's own limbs, src/core/ModSwitch.cu:
// 467-469): one 16-byte vector per thread, with non-temporal loads and
// stores since the data streams through once. For 1 GiB this reaches 6.6
// TB/s on an MI355X, where hipMemcpyAsync's copy kernel reaches 5.1 TB/s.
typedef uint32_t cp4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kThreads) void copy16_kernel(const cp4 *__restrict__ src,
                                                          cp4 *__restrict__ dst, size_t n) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

int launch_copy(const Context &ctx, void *dst, const void *src, size_t bytes,
                hipStream_t st) {
  if (bytes == 0 || dst == src) return kOk;
  const size_t n = bytes / 16, blocks = (n + kThreads - 1) / kThreads;
  const bool vec = (reinterpret_cast<uintptr_t>(dst) % 16) == 0 &&
                   (reinterpret_cast<uintptr_t>(src) % 16) == 0 && bytes % 16 == 0 &&
                   blocks <= 0x7fffffffu;
  if (!vec)  // unaligned or odd sizes: the runtime's copy
    return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st), "copy");
  launch(ctx, "copy", (int)std::min<size_t>(bytes / (65536 * 4), 1 << 30), copy16_kernel,
         dim3((unsigned)blocks), dim3(kThreads), st, static_cast<const cp4 *>(src),
         static_cast<cp4 *>(dst), n);
  return check_hip(hipGetLastError(), "copy");
}

static int bs_grid(int nlimbs) {
  return ((nlimbs + 7) / 8) * 8 * kBsBlocksPerLimb;
}

void launch_bs_fused(const Context &ctx, const EwNp &np, const BsFusedArgs &A,
                     hipStream_t st) {
  if (np.num_total <= 0 || A.nrot <= 0) return;
  launch(ctx, "bs_fused", np.num_total, bs_fused_kernel,
         dim3(bs_grid(np.num_total)), dim3(kThreads), st, np, ctx.ptab, A);
}

void launch_key_accum_table(const Context &ctx, const KeyAccumItem *items,
                            const EwJob *jobs, int njobs, hipStream_t st) {
  if (njobs <= 0) return;
  launch(ctx, "key_accum[batch]", njobs, key_accum_table_kernel,
         dim3(kBlocksPerLimb, njobs), dim3(kThreads), st, items, jobs,
         ctx.ptab);
}

void launch_bs_fused_table(const Context &ctx, const BsFusedItem *items,
                           const EwJob *jobs, int njobs, hipStream_t st) {
  if (njobs <= 0) return;
  launch(ctx, "bs_fused[batch]", njobs, bs_fused_table_kernel,
         dim3(bs_grid(njobs)), dim3(kThreads), st, items, jobs, njobs,
         ctx.ptab);
}

}  // namespace ck
