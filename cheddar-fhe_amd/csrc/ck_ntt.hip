// Negacyclic NTT / INTT over 32-bit RNS limbs for N = 2^16 on gfx950.
//
// Transform definition (the reference's, src/core/NTT.cu:231-454 and
// NTTUtils.cuh:19-58): forward Cooley-Tukey with bit-reversed powers of psi
// (psi = g^((q-1)/2N), g the smallest generator); stage m (= 2^s) pairs
// (j, j + N/2^(s+1)) in block i = j / (N/2^s) with twiddle psi_rev[m + i];
// output position k holds a(psi^(2 brv(k) + 1)). The inverse is the
// Gentleman-Sande mirror with psi^-1 and no N^-1 factor; the per-limb
// constant fused after it carries N^-1 and whatever else the caller folds.
//
// MI355X mapping. A limb of 2^16 u32 is a 256 x 256 matrix (row = 256
// consecutive coefficients).
//   * row kernels (forward stages 8..15, inverse stages 0..7): a 256-thread
//     workgroup owns 16 consecutive rows = one contiguous 16 KiB block,
//     moved with 16-byte loads/stores that cover whole cache lines.
//   * column kernels (forward stages 0..7, inverse 8..15): a 512-thread
//     workgroup owns all 256 rows x 32 columns (128 B per row segment).
// Each thread runs two radix-16 register phases on 16 values. Global data
// move through an int32 LDS staging tile with 16-byte accesses; the hand-off
// between the two phases keeps the lazy doubles in an LDS tile aliased with
// the staging tile (row tiles use an XOR swizzle so both phase layouts are
// free of bank conflicts). Twiddles are read once per workgroup phase as u32
// and issued before the data loads. Butterflies use the exact FP64 modular
// product of ck_arith.h with lazy ranges (integer doubles below 2^40); the
// intermediate between the two passes of a transform is stored as a lazy
// int32 in (-q/2 - 2, q/2 + 2) and only final outputs are made canonical.
// Each thread writes a phase's results back to exactly the LDS words it read
// for that phase, so only the hand-offs between threads need a barrier.
#include "ck_internal.h"

namespace ck {
namespace {

constexpr int kN = 65536;
constexpr int kRowThreads = 256;
constexpr int kRowTile = 16;                  // rows per row-kernel workgroup
constexpr int kRowTiles = 256 / kRowTile;     // per limb
constexpr int kColThreads = 512;
constexpr int kColTile = 32;                  // columns per column workgroup
constexpr int kColTiles = 256 / kColTile;     // per limb

// forward radix-16 (4 CT stages), element index bits high to low; twiddle
// slot (1 << s) - 1 + blk for sub-stage s
__device__ __forceinline__ void ct16(double x[16], const uint32_t tw[15],
                                     double q, double qr) {
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int half = 8 >> s;
#pragma unroll
    for (int blk = 0; blk < (1 << s); blk++) {
      const double w = (double)tw[(1 << s) - 1 + blk];
      const double wq = w * qr;
#pragma unroll
      for (int j = 0; j < half; j++) {
        double &a = x[blk * 2 * half + j];
        double &b = x[blk * 2 * half + j + half];
        const double r = mulmod_f(b, w, wq, q);
        b = a - r;
        a = a + r;
      }
    }
  }
}

// inverse radix-16 (4 GS stages), element index bits low to high; twiddle
// slot 16 - (16 >> u) + blk for sub-stage u
__device__ __forceinline__ void gs16(double x[16], const uint32_t tw[15],
                                     double q, double qr) {
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int half = 1 << u;
#pragma unroll
    for (int blk = 0; blk < (8 >> u); blk++) {
      const double w = (double)tw[16 - (16 >> u) + blk];
      const double wq = w * qr;
#pragma unroll
      for (int j = 0; j < half; j++) {
        double &a = x[blk * 2 * half + j];
        double &b = x[blk * 2 * half + j + half];
        const double d = a - b;
        a = a + b;
        b = mulmod_f(d, w, wq, q);
      }
    }
  }
}

// Row-tile LDS index: row rl, in-row offset b, XOR swizzle of the low nibble
// by the high nibble (both b = t + 16k and b = 16t + j patterns hit 16
// distinct banks).
__device__ __forceinline__ int rlds(int rl, int b) {
  return rl * 256 + (b ^ ((b >> 4) & 15));
}

// Launch descriptors: every kernel asks its descriptor for the LimbJob of
// limb y (uniform per workgroup). JobByValue unpacks the kernel-argument
// plan of a single operation; JobTable reads a batched plan's device table.
struct JobByValue {
  NttLaunch L;
  __device__ __forceinline__ LimbJob job(int y) const {
    const NttLimb &R = L.limb[y];
    const int sel = R.sel;
    const int b = (sel >> 6) & 3;
    LimbJob J;
    J.src = L.src[sel & 3] + (size_t)R.src_limb * kN;
    J.dst = L.dst[(sel >> 2) & 3] + (size_t)R.dst_limb * kN;
    const uint32_t *s2 = L.src2[(sel >> 4) & 3];
    J.src2 = (s2 && R.src2_limb >= 0) ? s2 + (size_t)R.src2_limb * kN : nullptr;
    J.bsrc = L.bsrc[b];
    J.btab = L.btab[b] ? L.btab[b] + R.btab_off : nullptr;
    J.bs = (uint16_t)L.bs[b];
    J.c1 = R.c1;
    J.c2 = R.c2;
    J.c3 = R.c3;
    J.prime = R.prime;
    J.flags = L.flags;
    return J;
  }
};

struct JobTable {
  const LimbJob *__restrict__ jobs;
  __device__ __forceinline__ LimbJob job(int y) const { return jobs[y]; }
};

// 16 KiB contiguous row tile <-> LDS, 16-byte accesses
__device__ __forceinline__ void tile_rows_to_lds(const uint32_t *g, int32_t *lds) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int e = 1024 * i + 4 * threadIdx.x;
    const uint4 v = *reinterpret_cast<const uint4 *>(g + e);
    const int rl = e >> 8, b = e & 255;
    lds[rlds(rl, b)] = (int32_t)v.x;
    lds[rlds(rl, b + 1)] = (int32_t)v.y;
    lds[rlds(rl, b + 2)] = (int32_t)v.z;
    lds[rlds(rl, b + 3)] = (int32_t)v.w;
  }
}

__device__ __forceinline__ void lds_to_tile_rows(const int32_t *lds, uint32_t *g) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int e = 1024 * i + 4 * threadIdx.x;
    const int rl = e >> 8, b = e & 255;
    uint4 v;
    v.x = (uint32_t)lds[rlds(rl, b)];
    v.y = (uint32_t)lds[rlds(rl, b + 1)];
    v.z = (uint32_t)lds[rlds(rl, b + 2)];
    v.w = (uint32_t)lds[rlds(rl, b + 3)];
    *reinterpret_cast<uint4 *>(g + e) = v;
  }
}

// column tile (256 rows x 32 columns starting at column c0) <-> LDS [256][32]
__device__ __forceinline__ void tile_cols_to_lds(const uint32_t *g, int c0,
                                                 int32_t *lds) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int idx = 512 * i + threadIdx.x;  // 16-byte chunk index
    const int r = idx >> 3, c4 = (idx & 7) * 4;
    const uint4 v = *reinterpret_cast<const uint4 *>(g + r * 256 + c0 + c4);
    *reinterpret_cast<int4 *>(lds + r * 32 + c4) =
        make_int4((int32_t)v.x, (int32_t)v.y, (int32_t)v.z, (int32_t)v.w);
  }
}

// BaseConv computed straight into the column tile (ModSwitchMatrixMult,
// src/core/ModSwitch.cu:23-171, in exact FP64): for each coefficient,
// y = sum_j v_j * T_j mod t over bs centred int32 source limbs, with
// T_j = hi_j * 2^16 + lo_j so every partial sum stays an exact double
// (|v| < 2^30: |sum v hi| < 2^49, |sum v lo| < 2^50 for bs <= 16).
__device__ __forceinline__ void bconv_cols_to_lds(const int32_t *bsrc, int bs,
                                                  const double *tab, int c0,
                                                  const ModF &m,
                                                  int32_t *lds) {
  double th[kMaxBconvSrc], tl[kMaxBconvSrc];
#pragma unroll
  for (int j = 0; j < kMaxBconvSrc; j++) {
    if (j < bs) {
      th[j] = tab[2 * j];
      tl[j] = tab[2 * j + 1];
    }
  }
  for (int i = 0; i < 4; i++) {
    const int idx = 512 * i + threadIdx.x;
    const int r = idx >> 3, c4 = (idx & 7) * 4;
    const int32_t *p = bsrc + r * 256 + c0 + c4;
    double sh[4] = {0.0, 0.0, 0.0, 0.0}, sl[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < kMaxBconvSrc; j++) {
      if (j < bs) {
        const int4 v = *reinterpret_cast<const int4 *>(p + (size_t)j * kN);
        const double vv[4] = {(double)v.x, (double)v.y, (double)v.z,
                              (double)v.w};
#pragma unroll
        for (int e = 0; e < 4; e++) {
          sh[e] = __builtin_fma(vv[e], th[j], sh[e]);
          sl[e] = __builtin_fma(vv[e], tl[j], sl[e]);
        }
      }
    }
    int4 o;
    int32_t *op = reinterpret_cast<int32_t *>(&o);
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const double big = sh[e] * 65536.0;  // exact (power-of-two scale)
      const double k = __builtin_rint(big * m.qinv);
      op[e] = (int32_t)reduce_f(__builtin_fma(-k, m.q, big) + sl[e], m);
    }
    *reinterpret_cast<int4 *>(lds + r * 32 + c4) = o;
  }
}

__device__ __forceinline__ void lds_to_tile_cols(const int32_t *lds, uint32_t *g,
                                                 int c0) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int idx = 512 * i + threadIdx.x;
    const int r = idx >> 3, c4 = (idx & 7) * 4;
    const int4 v = *reinterpret_cast<const int4 *>(lds + r * 32 + c4);
    *reinterpret_cast<uint4 *>(g + r * 256 + c0 + c4) =
        make_uint4((uint32_t)v.x, (uint32_t)v.y, (uint32_t)v.z, (uint32_t)v.w);
  }
}

// Final value of an element: canonical u32, or lazy int32 in
// (-q/2 - 2, q/2 + 2) for a pass-1 intermediate that pass 2 reads back.
__device__ __forceinline__ int32_t lazy_i32(double x, const ModF &m) {
  return (int32_t)reduce_f(x, m);
}

// ---------------------------------------------------------------------------
// Forward, column stages s = 0..7 -> lazy int32 intermediate
// ---------------------------------------------------------------------------
template <class D>
__global__ __launch_bounds__(kColThreads) void ntt_fwd_cols(
    const D desc, const PrimeTab P, const uint32_t *__restrict__ tw) {
  // int32 staging tile [256][32] aliased with the double exchange tile
  __shared__ double ldsd[256 * kColTile];
  int32_t *lds = reinterpret_cast<int32_t *>(ldsd);
  // XCD-aware: workgroups are dealt round-robin to the 8 XCDs, so column
  // tile (blockIdx.x & 7) of every limb runs on the same XCD and the source
  // tiles a BaseConv input shares stay in that XCD's L2.
  const int y = blockIdx.x >> 3;
  const int c0 = (blockIdx.x & 7) * kColTile;
  const LimbJob J = desc.job(y);
  const double q = P.qd[J.prime], qr = P.qr[J.prime];
  const ModF m{q, qr};
  const uint32_t *__restrict__ w = tw + (size_t)J.prime * kN;
  const int cc = threadIdx.x & 31;
  const int rb = threadIdx.x >> 5;  // 0..15

  // phase 1 twiddles psi_rev[2^s + blk] (uniform), phase 2 psi_rev[2^(4+s)
  // + rb 2^s + blk]
  uint32_t t1[15], t2[15];
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int blk = 0; blk < (1 << s); blk++) {
      t1[(1 << s) - 1 + blk] = w[(1 << s) + blk];
      t2[(1 << s) - 1 + blk] = w[(16 << s) + (rb << s) + blk];
    }
  if (J.flags & kNttBconvInput)
    bconv_cols_to_lds(J.bsrc, J.bs, J.btab, c0, m, lds);
  else
    tile_cols_to_lds(J.src, c0, lds);
  __syncthreads();

  double x[16];
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = (double)lds[(rb + 16 * k) * 32 + cc];
  if (J.flags & kNttPreMul) {
    const double cv = (double)J.c1, cq = cv * qr;
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = mulmod_f(x[k], cv, cq, q);
  }
  ct16(x, t1, q, qr);
  __syncthreads();  // staging words are reused by the double tile
#pragma unroll
  for (int k = 0; k < 16; k++) ldsd[(rb + 16 * k) * 32 + cc] = x[k];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; j++) x[j] = ldsd[(rb * 16 + j) * 32 + cc];
  ct16(x, t2, q, qr);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; j++) lds[(rb * 16 + j) * 32 + cc] = lazy_i32(x[j], m);
  __syncthreads();
  lds_to_tile_cols(lds, J.dst, c0);
}

// ---------------------------------------------------------------------------
// Forward, row stages s = 8..15 (+ optional ModDown epilogue)
// ---------------------------------------------------------------------------
template <class D>
__global__ __launch_bounds__(kRowThreads) void ntt_fwd_rows(
    const D desc, const PrimeTab P, const uint32_t *__restrict__ tw) {
  __shared__ double ldsd[kRowTile * 256];
  int32_t *lds = reinterpret_cast<int32_t *>(ldsd);
  const int y = blockIdx.y;
  const LimbJob J = desc.job(y);
  const double q = P.qd[J.prime], qr = P.qr[J.prime];
  const ModF m{q, qr};
  const uint32_t *__restrict__ w = tw + (size_t)J.prime * kN;
  const int rl = threadIdx.x >> 4;
  const int t16 = threadIdx.x & 15;
  const int r = blockIdx.x * kRowTile + rl;  // row within the limb
  const size_t tile = (size_t)blockIdx.x * kRowTile * 256;

  // phase 1 twiddles psi_rev[2^(8+s) + r 2^s + blk], phase 2
  // psi_rev[2^(12+s) + r 2^(4+s) + t16 2^s + blk]
  uint32_t t1[15], t2[15];
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int blk = 0; blk < (1 << s); blk++) {
      t1[(1 << s) - 1 + blk] = w[(256 << s) + (r << s) + blk];
      t2[(1 << s) - 1 + blk] =
          w[(4096 << s) + (r << (4 + s)) + (t16 << s) + blk];
    }
  // the pass-1 intermediate (in dst) is lazy signed int32
  tile_rows_to_lds(J.dst + tile, lds);
  __syncthreads();

  double x[16];
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = (double)lds[rlds(rl, t16 + 16 * k)];
  ct16(x, t1, q, qr);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) ldsd[rlds(rl, t16 + 16 * k)] = x[k];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; j++) x[j] = ldsd[rlds(rl, 16 * t16 + j)];
  ct16(x, t2, q, qr);

  // this thread now holds the 16 consecutive outputs b = 16 t16 + j of row r
  uint32_t out[16];
  if (J.flags & kNttModDownEpilogue) {
    // NTTPhase2ForModDown (NTT.cu:456-601): out = (src2 * pad - y) * ip
    const double ip = (double)J.c3, ipq = ip * qr;
    double a2[16];
    if (J.src2) {
      const uint32_t *__restrict__ sp = J.src2 + r * 256 + t16 * 16;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const uint4 u = reinterpret_cast<const uint4 *>(sp)[i];
        a2[4 * i + 0] = (double)u.x;
        a2[4 * i + 1] = (double)u.y;
        a2[4 * i + 2] = (double)u.z;
        a2[4 * i + 3] = (double)u.w;
      }
      if (J.flags & kNttEpiloguePad) {
        const double pv = (double)J.c2, pq = pv * qr;
#pragma unroll
        for (int j = 0; j < 16; j++) a2[j] = mulmod_f(a2[j], pv, pq, q);
      }
    } else {
#pragma unroll
      for (int j = 0; j < 16; j++) a2[j] = 0.0;
    }
#pragma unroll
    for (int j = 0; j < 16; j++)
      out[j] = canon_f(mulmod_f(a2[j] - x[j], ip, ipq, q), m);
  } else {
#pragma unroll
    for (int j = 0; j < 16; j++) out[j] = canon_f(x[j], m);
  }
  uint4 *o4 = reinterpret_cast<uint4 *>(J.dst + (size_t)r * 256 + t16 * 16);
#pragma unroll
  for (int i = 0; i < 4; i++)
    o4[i] = make_uint4(out[4 * i], out[4 * i + 1], out[4 * i + 2],
                       out[4 * i + 3]);
}

// ---------------------------------------------------------------------------
// Inverse, row stages u = 0..7 -> lazy int32 intermediate
// ---------------------------------------------------------------------------
template <class D>
__global__ __launch_bounds__(kRowThreads) void intt_rows(
    const D desc, const PrimeTab P, const uint32_t *__restrict__ tw) {
  __shared__ double ldsd[kRowTile * 256];
  int32_t *lds = reinterpret_cast<int32_t *>(ldsd);
  const int y = blockIdx.y;
  const LimbJob J = desc.job(y);
  const double q = P.qd[J.prime], qr = P.qr[J.prime];
  const ModF m{q, qr};
  const uint32_t *__restrict__ w = tw + (size_t)J.prime * kN;
  const int rl = threadIdx.x >> 4;
  const int t16 = threadIdx.x & 15;
  const int r = blockIdx.x * kRowTile + rl;
  const size_t tile = (size_t)blockIdx.x * kRowTile * 256;

  // phase 1 psi_inv_rev[2^(15-u) + r 2^(7-u) + t16 2^(3-u) + blk],
  // phase 2 psi_inv_rev[2^(11-u) + r 2^(3-u) + blk]
  uint32_t t1[15], t2[15];
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int blk = 0; blk < (8 >> u); blk++) {
      t1[16 - (16 >> u) + blk] =
          w[(32768 >> u) + (r << (7 - u)) + (t16 << (3 - u)) + blk];
      t2[16 - (16 >> u) + blk] = w[(2048 >> u) + (r << (3 - u)) + blk];
    }
  // phase 1 reads b = 16 t16 + j: 64 contiguous bytes per thread
  double x[16];
  {
    const uint4 *s4 =
        reinterpret_cast<const uint4 *>(J.src + (size_t)r * 256 + t16 * 16);
    if (J.flags & kInttLoadProduct) {
      // Tensor's a1 (*) a2 (ElementWise.cu:504-546) fused into the load
      const uint32_t qu = P.q[J.prime];
      const int32_t qinv = P.qinv[J.prime];
      const uint4 *t4 = reinterpret_cast<const uint4 *>(
          J.src2 + (size_t)r * 256 + t16 * 16);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const uint4 u = s4[i], w2 = t4[i];
        x[4 * i + 0] = (double)mont_mul(u.x, w2.x, qu, qinv);
        x[4 * i + 1] = (double)mont_mul(u.y, w2.y, qu, qinv);
        x[4 * i + 2] = (double)mont_mul(u.z, w2.z, qu, qinv);
        x[4 * i + 3] = (double)mont_mul(u.w, w2.w, qu, qinv);
      }
    } else {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const uint4 u = s4[i];
        x[4 * i + 0] = (double)u.x;
        x[4 * i + 1] = (double)u.y;
        x[4 * i + 2] = (double)u.z;
        x[4 * i + 3] = (double)u.w;
      }
    }
  }
  gs16(x, t1, q, qr);
  // GS sums double every stage: reduce before the next four stages
#pragma unroll
  for (int j = 0; j < 16; j++) ldsd[rlds(rl, 16 * t16 + j)] = reduce_f(x[j], m);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = ldsd[rlds(rl, t16 + 16 * k)];
  gs16(x, t2, q, qr);
  __syncthreads();  // int32 words alias doubles other threads still read
#pragma unroll
  for (int k = 0; k < 16; k++) lds[rlds(rl, t16 + 16 * k)] = lazy_i32(x[k], m);
  __syncthreads();
  lds_to_tile_rows(lds, J.dst + tile);
}

// ---------------------------------------------------------------------------
// Inverse, column stages u = 8..15, then the per-limb constant
// ---------------------------------------------------------------------------
template <class D>
__global__ __launch_bounds__(kColThreads) void intt_cols(
    const D desc, const PrimeTab P, const uint32_t *__restrict__ tw) {
  __shared__ double ldsd[256 * kColTile];
  int32_t *lds = reinterpret_cast<int32_t *>(ldsd);
  const int y = blockIdx.x >> 3;  // same XCD-aware order as ntt_fwd_cols
  const int c0 = (blockIdx.x & 7) * kColTile;
  const LimbJob J = desc.job(y);
  const double q = P.qd[J.prime], qr = P.qr[J.prime];
  const ModF m{q, qr};
  const uint32_t *__restrict__ w = tw + (size_t)J.prime * kN;
  const int cc = threadIdx.x & 31;
  const int rb = threadIdx.x >> 5;

  // phase 1 psi_inv_rev[2^(7-u) + rb 2^(3-u) + blk],
  // phase 2 psi_inv_rev[2^(3-u) + blk] (uniform)
  uint32_t t1[15], t2[15];
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int blk = 0; blk < (8 >> u); blk++) {
      t1[16 - (16 >> u) + blk] = w[(128 >> u) + (rb << (3 - u)) + blk];
      t2[16 - (16 >> u) + blk] = w[(8 >> u) + blk];
    }
  // the pass-1 intermediate (in dst) is lazy signed int32
  tile_cols_to_lds(J.dst, c0, lds);
  __syncthreads();
  double x[16];
#pragma unroll
  for (int j = 0; j < 16; j++) x[j] = (double)lds[(rb * 16 + j) * 32 + cc];
  gs16(x, t1, q, qr);
  __syncthreads();  // staging words are reused by the double tile
#pragma unroll
  for (int j = 0; j < 16; j++)
    ldsd[(rb * 16 + j) * 32 + cc] = reduce_f(x[j], m);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = ldsd[(rb + 16 * k) * 32 + cc];
  gs16(x, t2, q, qr);
  const double cv = (double)J.c1, cq = cv * qr;
  const bool centred = (J.flags & kInttCentred) != 0;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const double r0 = reduce_f(mulmod_f(x[k], cv, cq, q), m);
    // canonical [0, q) or centred [-(q-1)/2, (q-1)/2]
    const double lo = centred ? -(q - 1.0) * 0.5 : 0.0;
    double r1 = r0 < lo ? r0 + q : r0;
    r1 = r1 >= lo + q ? r1 - q : r1;
    lds[(rb + 16 * k) * 32 + cc] = (int32_t)r1;
  }
  __syncthreads();
  lds_to_tile_cols(lds, J.dst, c0);
}

const char *fwd_cols_name(int flags) {
  return (flags & kNttBconvInput) ? "ntt_fwd_cols+bconv" : "ntt_fwd_cols";
}
const char *fwd_rows_name(int flags) {
  return (flags & kNttModDownEpilogue) ? "ntt_fwd_rows+epilogue" : "ntt_fwd_rows";
}
const char *inv_rows_name(int flags) {
  return (flags & kInttLoadProduct) ? "intt_rows+tensor" : "intt_rows";
}

template <class D>
void fwd_passes(const Context &ctx, const D &d, int n, int flags,
                hipStream_t st) {
  launch(ctx, fwd_cols_name(flags), n, ntt_fwd_cols<D>, dim3(kColTiles * n),
         dim3(kColThreads), st, d, ctx.ptab, (const uint32_t *)ctx.d_tw_fwd);
  launch(ctx, fwd_rows_name(flags), n, ntt_fwd_rows<D>, dim3(kRowTiles, n),
         dim3(kRowThreads), st, d, ctx.ptab, (const uint32_t *)ctx.d_tw_fwd);
}

template <class D>
void inv_passes(const Context &ctx, const D &d, int n, int flags,
                hipStream_t st) {
  launch(ctx, inv_rows_name(flags), n, intt_rows<D>, dim3(kRowTiles, n),
         dim3(kRowThreads), st, d, ctx.ptab, (const uint32_t *)ctx.d_tw_inv);
  launch(ctx, "intt_cols", n, intt_cols<D>, dim3(kColTiles * n),
         dim3(kColThreads), st, d, ctx.ptab, (const uint32_t *)ctx.d_tw_inv);
}

}  // namespace

// Forward NTT of L.n limbs: column stages into dst, then row stages in place
// on dst (with the optional ModDown epilogue).
void launch_ntt_fwd(const Context &ctx, const NttLaunch &L, hipStream_t st) {
  if (L.n <= 0) return;
  fwd_passes(ctx, JobByValue{L}, L.n, L.flags, st);
}

// Inverse NTT of L.n limbs: row stages into dst, then column stages in
// place on dst with the post constant c1 (plain form) and optional centring.
void launch_intt(const Context &ctx, const NttLaunch &L, hipStream_t st) {
  if (L.n <= 0) return;
  inv_passes(ctx, JobByValue{L}, L.n, L.flags, st);
}

void launch_ntt_fwd_table(const Context &ctx, const LimbJob *jobs, int n,
                          int flags_any, hipStream_t st) {
  if (n <= 0) return;
  fwd_passes(ctx, JobTable{jobs}, n, flags_any, st);
}

void launch_intt_table(const Context &ctx, const LimbJob *jobs, int n,
                       int flags_any, hipStream_t st) {
  if (n <= 0) return;
  inv_passes(ctx, JobTable{jobs}, n, flags_any, st);
}

}  // namespace ck
