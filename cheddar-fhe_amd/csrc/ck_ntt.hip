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
// MI355X mapping. A limb of 2^16 u32 is viewed as 256 rows x 256 columns.
//   * "column" kernels run the 8 stages whose butterflies pair rows
//     (s = 0..7 forward, u = 8..15 inverse). A 256-thread workgroup owns a
//     tile of 256 rows x 16 columns; each thread holds 16 values of one
//     column and runs two radix-16 register phases separated by one LDS
//     transpose.
//   * "row" kernels run the 8 stages inside a 256-element row. A
//     workgroup owns 16 rows; 16 threads per row, 16 values per thread.
// Every butterfly uses the exact FP64 modular product of ck_arith.h (full
// rate FMA on CDNA4) with lazy ranges: values stay integer doubles well
// below 2^48 inside a kernel and are reduced to canonical u32 on store, so
// HBM traffic stays at 4 B per coefficient per pass.
#include "ck_internal.h"

namespace ck {
namespace {

constexpr int kLogN = 16;
constexpr int kN = 1 << kLogN;
constexpr int kThreads = 256;
constexpr int kTile = 16;  // columns per column-tile, rows per row-tile
constexpr int kTilesPerLimb = 256 / kTile;
constexpr int kLdsStride = 256 + 16;  // padded row of the row kernels
constexpr int kColStride = kTile + 1; // padded row of the column kernels

struct Tw16 {
  double w[15];
  double wq[15];
};

__device__ __forceinline__ void set_tw(Tw16 &t, int i, uint32_t w, double qr) {
  double d = (double)w;
  t.w[i] = d;
  t.wq[i] = d * qr;
}

// forward radix-16 (4 CT stages), element index bits high to low;
// twiddle slot (1 << s) - 1 + blk for sub-stage s
__device__ __forceinline__ void ct16(double x[16], const Tw16 &t, double q) {
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int half = 8 >> s;
#pragma unroll
    for (int blk = 0; blk < (1 << s); blk++) {
      const double w = t.w[(1 << s) - 1 + blk];
      const double wq = t.wq[(1 << s) - 1 + blk];
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

// inverse radix-16 (4 GS stages), element index bits low to high;
// twiddle slot 16 - (16 >> u) + blk for sub-stage u
__device__ __forceinline__ void gs16(double x[16], const Tw16 &t, double q) {
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int half = 1 << u;
#pragma unroll
    for (int blk = 0; blk < (8 >> u); blk++) {
      const double w = t.w[16 - (16 >> u) + blk];
      const double wq = t.wq[16 - (16 >> u) + blk];
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

// ---------------------------------------------------------------------------
// Forward, column stages s = 0..7
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void ntt_fwd_cols(
    const NttLaunch L, const PrimeTab P, const uint32_t *__restrict__ tw) {
  __shared__ double lds[256 * kColStride];
  const int y = blockIdx.y;
  const int p = L.prime[y];
  const double q = P.qd[p], qr = P.qr[p];
  const ModF m{q, qr};
  const uint32_t *__restrict__ src = L.src[L.sel[y] & 3] + (size_t)L.src_limb[y] * kN;
  uint32_t *__restrict__ dst = L.dst[(L.sel[y] >> 2) & 3] + (size_t)L.dst_limb[y] * kN;
  const uint32_t *__restrict__ w = tw + (size_t)p * kN;
  const int cc = threadIdx.x & (kTile - 1);
  const int rb = threadIdx.x >> 4;
  const int c = blockIdx.x * kTile + cc;

  double x[16];
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = (double)src[(rb + 16 * k) * 256 + c];
  if (L.flags & kNttPreMul) {
    const double cv = (double)L.c1[y], cq = cv * qr;
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = mulmod_f(x[k], cv, cq, q);
  }
  Tw16 t;
  // phase 1: rows rb + 16k, stages 0..3, twiddle psi_rev[2^s + blk]
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int blk = 0; blk < (1 << s); blk++)
      set_tw(t, (1 << s) - 1 + blk, w[(1 << s) + blk], qr);
  ct16(x, t, q);
#pragma unroll
  for (int k = 0; k < 16; k++) lds[(rb + 16 * k) * kColStride + cc] = x[k];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; j++) x[j] = lds[(rb * 16 + j) * kColStride + cc];
  // phase 2: rows rb*16 + j, stages 4..7,
  // twiddle psi_rev[2^(4+s) + rb*2^s + blk]
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int blk = 0; blk < (1 << s); blk++)
      set_tw(t, (1 << s) - 1 + blk, w[(16 << s) + (rb << s) + blk], qr);
  ct16(x, t, q);
#pragma unroll
  for (int j = 0; j < 16; j++) dst[(rb * 16 + j) * 256 + c] = canon_f(x[j], m);
}

// ---------------------------------------------------------------------------
// Forward, row stages s = 8..15 (+ optional ModDown epilogue)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void ntt_fwd_rows(
    const NttLaunch L, const PrimeTab P, const uint32_t *__restrict__ tw) {
  __shared__ double lds[kTile * kLdsStride];
  const int y = blockIdx.y;
  const int p = L.prime[y];
  const double q = P.qd[p], qr = P.qr[p];
  const ModF m{q, qr};
  const int rl = threadIdx.x >> 4;  // row within tile
  const int t16 = threadIdx.x & 15;
  const int r = blockIdx.x * kTile + rl;  // row within limb
  const uint32_t *__restrict__ src =
      L.src[L.sel[y] & 3] + (size_t)L.src_limb[y] * kN + r * 256;
  uint32_t *__restrict__ dst = L.dst[(L.sel[y] >> 2) & 3] + (size_t)L.dst_limb[y] * kN + r * 256;
  const uint32_t *__restrict__ w = tw + (size_t)p * kN;
  double *row = lds + rl * kLdsStride;

  double x[16];
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = (double)src[t16 + 16 * k];
  Tw16 t;
  // phase 1: b = t16 + 16k, stages 8..11, twiddle psi_rev[2^(8+s) + r*2^s +
  // blk]
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int blk = 0; blk < (1 << s); blk++)
      set_tw(t, (1 << s) - 1 + blk, w[(256 << s) + (r << s) + blk], qr);
  ct16(x, t, q);
#pragma unroll
  for (int k = 0; k < 16; k++) row[t16 + 17 * k] = x[k];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; j++) x[j] = row[17 * t16 + j];
  // phase 2: b = t16*16 + j, stages 12..15,
  // twiddle psi_rev[2^(12+s) + r*2^(4+s) + t16*2^s + blk]
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int blk = 0; blk < (1 << s); blk++)
      set_tw(t, (1 << s) - 1 + blk,
             w[(4096 << s) + (r << (4 + s)) + (t16 << s) + blk], qr);
  ct16(x, t, q);

  uint32_t out[16];
  if (L.flags & kNttModDownEpilogue) {
    // NTTPhase2ForModDown (NTT.cu:456-601): out = (src2 * pad - y) * ip
    const int s2 = L.src2_limb[y];
    const double ip = (double)L.c3[y], ipq = ip * qr;
    double a2[16];
    if (s2 >= 0) {
      const uint32_t *__restrict__ sp =
          L.src2[(L.sel[y] >> 4) & 3] + (size_t)s2 * kN + r * 256 + t16 * 16;
      const uint4 *v4 = reinterpret_cast<const uint4 *>(sp);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        uint4 v = v4[i];
        a2[4 * i + 0] = (double)v.x;
        a2[4 * i + 1] = (double)v.y;
        a2[4 * i + 2] = (double)v.z;
        a2[4 * i + 3] = (double)v.w;
      }
      if (L.flags & kNttEpiloguePad) {
        const double pv = (double)L.c2[y], pq = pv * qr;
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
  uint4 *o4 = reinterpret_cast<uint4 *>(dst + t16 * 16);
#pragma unroll
  for (int i = 0; i < 4; i++)
    o4[i] = make_uint4(out[4 * i], out[4 * i + 1], out[4 * i + 2],
                       out[4 * i + 3]);
}

// ---------------------------------------------------------------------------
// Inverse, row stages u = 0..7
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void intt_rows(
    const NttLaunch L, const PrimeTab P, const uint32_t *__restrict__ tw) {
  __shared__ double lds[kTile * kLdsStride];
  const int y = blockIdx.y;
  const int p = L.prime[y];
  const double q = P.qd[p], qr = P.qr[p];
  const ModF m{q, qr};
  const int rl = threadIdx.x >> 4;
  const int t16 = threadIdx.x & 15;
  const int r = blockIdx.x * kTile + rl;
  const uint32_t *__restrict__ src =
      L.src[L.sel[y] & 3] + (size_t)L.src_limb[y] * kN + r * 256;
  uint32_t *__restrict__ dst = L.dst[(L.sel[y] >> 2) & 3] + (size_t)L.dst_limb[y] * kN + r * 256;
  const uint32_t *__restrict__ w = tw + (size_t)p * kN;
  double *row = lds + rl * kLdsStride;

  double x[16];
  {
    const uint4 *v4 = reinterpret_cast<const uint4 *>(src + t16 * 16);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      uint4 v = v4[i];
      x[4 * i + 0] = (double)v.x;
      x[4 * i + 1] = (double)v.y;
      x[4 * i + 2] = (double)v.z;
      x[4 * i + 3] = (double)v.w;
    }
  }
  Tw16 t;
  // phase 1: b = t16*16 + j, stages u = 0..3,
  // twiddle psi_inv_rev[2^(15-u) + r*2^(7-u) + t16*2^(3-u) + blk]
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int blk = 0; blk < (8 >> u); blk++)
      set_tw(t, 16 - (16 >> u) + blk,
             w[(32768 >> u) + (r << (7 - u)) + (t16 << (3 - u)) + blk], qr);
  gs16(x, t, q);
#pragma unroll
  for (int j = 0; j < 16; j++) row[17 * t16 + j] = x[j];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = row[t16 + 17 * k];
  // phase 2: b = t16 + 16k, stages u = 4..7,
  // twiddle psi_inv_rev[2^(11-u') + r*2^(3-u') + blk]
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int blk = 0; blk < (8 >> u); blk++)
      set_tw(t, 16 - (16 >> u) + blk,
             w[(2048 >> u) + (r << (3 - u)) + blk], qr);
  gs16(x, t, q);
#pragma unroll
  for (int k = 0; k < 16; k++) dst[t16 + 16 * k] = canon_f(x[k], m);
}

// ---------------------------------------------------------------------------
// Inverse, column stages u = 8..15, then the per-limb constant
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void intt_cols(
    const NttLaunch L, const PrimeTab P, const uint32_t *__restrict__ tw) {
  __shared__ double lds[256 * kColStride];
  const int y = blockIdx.y;
  const int p = L.prime[y];
  const double q = P.qd[p], qr = P.qr[p];
  const ModF m{q, qr};
  const uint32_t *__restrict__ src = L.src[L.sel[y] & 3] + (size_t)L.src_limb[y] * kN;
  uint32_t *__restrict__ dst = L.dst[(L.sel[y] >> 2) & 3] + (size_t)L.dst_limb[y] * kN;
  const uint32_t *__restrict__ w = tw + (size_t)p * kN;
  const int cc = threadIdx.x & (kTile - 1);
  const int rb = threadIdx.x >> 4;
  const int c = blockIdx.x * kTile + cc;

  double x[16];
#pragma unroll
  for (int j = 0; j < 16; j++) x[j] = (double)src[(rb * 16 + j) * 256 + c];
  Tw16 t;
  // phase 1: rows rb*16 + j, stages u = 8..11,
  // twiddle psi_inv_rev[2^(7-u') + rb*2^(3-u') + blk]
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int blk = 0; blk < (8 >> u); blk++)
      set_tw(t, 16 - (16 >> u) + blk, w[(128 >> u) + (rb << (3 - u)) + blk],
             qr);
  gs16(x, t, q);
#pragma unroll
  for (int j = 0; j < 16; j++) lds[(rb * 16 + j) * kColStride + cc] = x[j];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = lds[(rb + 16 * k) * kColStride + cc];
  // phase 2: rows rb + 16k, stages u = 12..15, twiddle psi_inv_rev[2^(3-u') +
  // blk]
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int blk = 0; blk < (8 >> u); blk++)
      set_tw(t, 16 - (16 >> u) + blk, w[(8 >> u) + blk], qr);
  gs16(x, t, q);
  const double cv = (double)L.c1[y], cq = cv * qr;
  if (L.flags & kInttCentred) {
#pragma unroll
    for (int k = 0; k < 16; k++)
      dst[(rb + 16 * k) * 256 + c] =
          (uint32_t)centre_f(mulmod_f(x[k], cv, cq, q), m);
  } else {
#pragma unroll
    for (int k = 0; k < 16; k++)
      dst[(rb + 16 * k) * 256 + c] = canon_f(mulmod_f(x[k], cv, cq, q), m);
  }
}

// second pass reads what the first pass wrote
void in_place(NttLaunch &L) {
  for (int i = 0; i < 4; i++) L.src[i] = L.dst[i];
  for (int i = 0; i < L.n; i++) {
    L.src_limb[i] = L.dst_limb[i];
    L.sel[i] = (uint8_t)((L.sel[i] & ~3) | ((L.sel[i] >> 2) & 3));
  }
}

}  // namespace

// Forward NTT of L.n limbs: column stages into dst, then row stages in place
// on dst (with the optional ModDown epilogue).
void launch_ntt_fwd(const Context &ctx, const NttLaunch &L, hipStream_t st) {
  if (L.n <= 0) return;
  dim3 grid(kTilesPerLimb, L.n);
  hipLaunchKernelGGL(ntt_fwd_cols, grid, dim3(kThreads), 0, st, L, ctx.ptab,
                     ctx.d_tw_fwd);
  NttLaunch L2 = L;
  in_place(L2);
  hipLaunchKernelGGL(ntt_fwd_rows, grid, dim3(kThreads), 0, st, L2, ctx.ptab,
                     ctx.d_tw_fwd);
}

// Inverse NTT of L.n limbs: row stages into dst, then column stages in
// place on dst with the post constant c1 (plain form) and optional centring.
void launch_intt(const Context &ctx, const NttLaunch &L, hipStream_t st) {
  if (L.n <= 0) return;
  dim3 grid(kTilesPerLimb, L.n);
  hipLaunchKernelGGL(intt_rows, grid, dim3(kThreads), 0, st, L, ctx.ptab,
                     ctx.d_tw_inv);
  NttLaunch L2 = L;
  in_place(L2);
  hipLaunchKernelGGL(intt_cols, grid, dim3(kThreads), 0, st, L2, ctx.ptab,
                     ctx.d_tw_inv);
}

}  // namespace ck
