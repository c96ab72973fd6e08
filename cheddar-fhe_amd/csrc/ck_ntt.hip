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
//     workgroup owns 16 consecutive rows = one contiguous 16 KiB block; the
//     side a thread holds 16 consecutive words of moves with 16-byte
//     accesses, the other side straight between registers and memory with
//     4-byte accesses (64 contiguous bytes per row and instruction).
//   * column kernels (forward stages 0..7, inverse 8..15): a 256-thread
//     workgroup owns all 256 rows x 16 columns; every load and store is a
//     64-byte row segment per 16 lanes, read and written straight from
//     registers (32 KiB of LDS: four workgroups and more fit a CU).
// Each thread runs two radix-16 register phases on 16 values; the hand-off
// between the two phases keeps the lazy doubles in an LDS tile (row tiles
// use an XOR swizzle so both phase layouts are free of bank conflicts; the
// column pass's BaseConv stages its int32 output in the same LDS words
// first). Twiddles are read once per workgroup phase as u32
// and issued before the data loads. Butterflies use the exact FP64 modular
// product of ck_arith.h with lazy ranges (integer doubles below 2^40); the
// intermediate between the two passes of a transform is stored as a lazy
// int32 in (-0.75 q, 0.75 q) and only final outputs are made canonical.
// Each thread writes a phase's results back to exactly the LDS words it read
// for that phase, so only the hand-offs between threads need a barrier.
#include <algorithm>

#include "ck_elem.h"
#include "ck_internal.h"

namespace ck {
namespace {

constexpr int kN = 65536;
constexpr int kRowThreads = 256;
constexpr int kRowTile = 16;                  // rows per row-kernel workgroup
constexpr int kRowTiles = 256 / kRowTile;     // per limb
constexpr int kColThreads = 256;
constexpr int kColTile = 16;                  // columns per column workgroup
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
#pragma unroll
      for (int j = 0; j < half; j++) {
        double &a = x[blk * 2 * half + j];
        double &b = x[blk * 2 * half + j + half];
        const double r = mulmod_fq(b, w, q, qr);
        b = a - r;
        a = a + r;
      }
    }
  }
}

// ct16 with the twiddles kept in LDS, slot i of this thread at
// tws[i * kRowThreads + threadIdx.x] (conflict-free): for a loop that runs
// several transforms with the same twiddles and cannot keep them, and their
// FP64 forms, in registers.
__device__ __forceinline__ void ct16_lds(double x[16], const uint32_t *tws,
                                         double q, double qr) {
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int half = 8 >> s;
#pragma unroll
    for (int blk = 0; blk < (1 << s); blk++) {
      const double w = (double)tws[((1 << s) - 1 + blk) * 256 + threadIdx.x];
#pragma unroll
      for (int j = 0; j < half; j++) {
        double &a = x[blk * 2 * half + j];
        double &b = x[blk * 2 * half + j + half];
        const double r = mulmod_fq(b, w, q, qr);
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
#pragma unroll
      for (int j = 0; j < half; j++) {
        double &a = x[blk * 2 * half + j];
        double &b = x[blk * 2 * half + j + half];
        const double d = a - b;
        a = a + b;
        b = mulmod_fq(d, w, q, qr);
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
  // the record is read-only and uniform per workgroup: scalar loads
  // through the constant address space
  __device__ __forceinline__ LimbJob job(int y) const {
    typedef const __attribute__((address_space(4))) LimbJob CJob;
    CJob *p = (CJob *)jobs + y;
    LimbJob J;
    J.src = p->src;
    J.dst = p->dst;
    J.src2 = p->src2;
    J.bsrc = p->bsrc;
    J.btab = p->btab;
    J.c1 = p->c1;
    J.c2 = p->c2;
    J.c3 = p->c3;
    J.prime = p->prime;
    J.bs = p->bs;
    J.flags = p->flags;
    return J;
  }
};

// Row tile element b = t16 + 16 k of row rl straight from / to global
// memory: each load or store instruction covers 64 contiguous bytes of each
// of the wave's four rows, and no LDS staging or barrier is needed.
__device__ __forceinline__ void rows_load_p1(const uint32_t *tile, int rl, int t16,
                                             double x[16]) {
  typedef const __attribute__((address_space(1))) int32_t gint;
  gint *g = (gint *)(tile + rl * 256 + t16);
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = (double)g[16 * k];
}
__device__ __forceinline__ void rows_store_p1(uint32_t *tile, int rl, int t16,
                                              const int32_t v[16]) {
  typedef __attribute__((address_space(1))) int32_t gint;
  gint *g = (gint *)(tile + rl * 256 + t16);
#pragma unroll
  for (int k = 0; k < 16; k++) g[16 * k] = v[k];
}

// Column tile of a launch: 16 tiles of 16 columns per limb; workgroup i
// runs on XCD i mod 8, and tiles 2x and 2x + 1 (which share 128-byte lines)
// are placed on the same XCD so a line is fetched into one L2 only.
struct ColTile {
  int y, c0;
};
__device__ __forceinline__ ColTile col_tile() {
  const int t = ((blockIdx.x & 7) << 1) | ((blockIdx.x >> 3) & 1);
  return ColTile{(int)(blockIdx.x >> 4), t * kColTile};
}

// BaseConv computed straight into the column tile (ModSwitchMatrixMult,
// src/core/ModSwitch.cu:23-171, in exact FP64): for each coefficient,
// y = sum_j v_j * T_j mod t over bs centred int32 source limbs, with
// T_j = hi_j * 2^16 + lo_j so every partial sum stays an exact double
// (|v| < 2^30, hi < 2^15, lo < 2^16: |sum v hi| < 2^50 and
// |sum v lo| < 2^51 for bs <= 24).
// Output: int32 tile [256][16] in LDS, 16-byte source loads.
//
// MS (>= bs) bounds the unrolled source loop. All loads
// of a batch of up to KB sources are issued unconditionally before the first
// product (sources past bs re-read source 0 and get zero table entries): a
// load inside a branch that also consumes it would be waited for at once,
// one memory latency per source. VEC coefficients per load: 4 (16 bytes),
// or 2 for MS >= 12, whose 24 or 32 table doubles leave fewer registers.
template <int MS, int VEC, int KB = 8>
__device__ __forceinline__ void bconv_cols_ms(const int32_t *bsrc, int bs,
                                              const double *tab, int c0,
                                              const ModF &m, int32_t *lds) {
  typedef int32_t ivec __attribute__((ext_vector_type(VEC)));
  constexpr int kRowLanes = kColTile / VEC;                 // lanes per row
  constexpr int kIters = 256 * kColTile / (kColThreads * VEC);
  // the table row is read-only and uniform: scalar loads (constant space)
  typedef const __attribute__((address_space(4))) double cdouble;
  cdouble *ctab = (cdouble *)tab;
  double th[MS], tl[MS];
#pragma unroll
  for (int j = 0; j < MS; j++) {
    const int jj = j < bs ? j : 0;
    const double h = ctab[2 * jj], l = ctab[2 * jj + 1];
    th[j] = j < bs ? h : 0.0;
    tl[j] = j < bs ? l : 0.0;
  }
#pragma unroll
  for (int i = 0; i < kIters; i++) {
    const int idx = kColThreads * i + threadIdx.x;  // VEC-word chunk index
    const int r = idx / kRowLanes, cv = (idx % kRowLanes) * VEC;
    const int32_t *p = bsrc + r * 256 + c0 + cv;
    double sh[VEC], sl[VEC];
#pragma unroll
    for (int e = 0; e < VEC; e++) sh[e] = sl[e] = 0.0;
#pragma unroll
    for (int j0 = 0; j0 < MS; j0 += KB) {
      constexpr int kB = MS < KB ? MS : KB;
      ivec v[kB];
#pragma unroll
      for (int j = 0; j < kB; j++) {
        if (j0 + j >= MS) break;  // static after unrolling
        const int jj = j0 + j < bs ? j0 + j : 0;
        v[j] = *reinterpret_cast<const ivec *>(p + (size_t)jj * kN);
      }
#pragma unroll
      for (int j = 0; j < kB; j++) {
        if (j0 + j >= MS) break;
#pragma unroll
        for (int e = 0; e < VEC; e++) {
          const double vv = (double)v[j][e];
          sh[e] = __builtin_fma(vv, th[j0 + j], sh[e]);
          sl[e] = __builtin_fma(vv, tl[j0 + j], sl[e]);
        }
      }
    }
    ivec o;
#pragma unroll
    for (int e = 0; e < VEC; e++) {
      const double big = sh[e] * 65536.0;  // exact (power-of-two scale)
      const double k = __builtin_rint(big * m.qinv);
      o[e] = (int32_t)reduce_f(__builtin_fma(-k, m.q, big) + sl[e], m);
    }
    *reinterpret_cast<ivec *>(lds + r * kColTile + cv) = o;
  }
}

// bs is uniform per workgroup: one branch-free body per source bound.
// WIDTH 1: the instantiation for launches with 10..16 sources, whose
// 12- and 13-source bodies keep 16-byte loads at four waves per SIMD (at
// five they only fit with 8-byte loads); WIDTH 2: 17..24 sources (terminal
// plus auxiliary primes of a rescale), at three waves per SIMD
template <int WIDTH>
__device__ __forceinline__ void bconv_cols_to_lds(const int32_t *bsrc, int bs,
                                                  const double *tab, int c0,
                                                  const ModF &m,
                                                  int32_t *lds) {
  // exact bodies for the common counts (alpha and 1 + alpha of the shipped
  // parameter sets: 8, 9, 12, 13), bounded bodies otherwise
  switch (bs) {
    case 8: bconv_cols_ms<8, 4>(bsrc, bs, tab, c0, m, lds); return;
    case 9: bconv_cols_ms<9, 4, 5>(bsrc, bs, tab, c0, m, lds); return;
    case 12:
      if (WIDTH >= 1)
        bconv_cols_ms<12, 4, 6>(bsrc, bs, tab, c0, m, lds);
      else
        bconv_cols_ms<12, 2>(bsrc, bs, tab, c0, m, lds);
      return;
    case 13:
      if (WIDTH >= 1)
        bconv_cols_ms<13, 4, 5>(bsrc, bs, tab, c0, m, lds);
      else
        bconv_cols_ms<13, 2>(bsrc, bs, tab, c0, m, lds);
      return;
    default: break;
  }
  if (bs <= 4)
    bconv_cols_ms<4, 4>(bsrc, bs, tab, c0, m, lds);
  else if (bs <= 8)
    bconv_cols_ms<8, 4>(bsrc, bs, tab, c0, m, lds);
  else if (bs <= 12)
    bconv_cols_ms<12, 2>(bsrc, bs, tab, c0, m, lds);
  else if (WIDTH < 2 || bs <= 16)
    bconv_cols_ms<16, 2>(bsrc, bs, tab, c0, m, lds);
  else if constexpr (WIDTH == 2)
    bconv_cols_ms<kMaxBconvSrc, 2, 6>(bsrc, bs, tab, c0, m, lds);
}

// Final value of an element: canonical u32, or lazy int32 in
// (-0.75 q, 0.75 q) for a pass-1 intermediate that pass 2 reads back.
__device__ __forceinline__ int32_t lazy_i32(double x, const ModF &m) {
  return (int32_t)reduce_f(x, m);
}

// ---------------------------------------------------------------------------
// Forward, column stages s = 0..7 -> lazy int32 intermediate
// ---------------------------------------------------------------------------
template <class D, int WIDTH>
__global__ __launch_bounds__(kColThreads, 5 - WIDTH) void ntt_fwd_cols(
    const D desc, const PrimeTab P, const uint32_t *__restrict__ tw) {
  // int32 BaseConv tile [256][16] aliased with the double exchange tile
  __shared__ double ldsd[256 * kColTile];
  int32_t *lds = reinterpret_cast<int32_t *>(ldsd);
  const ColTile ct = col_tile();
  const int c0 = ct.c0;
  const LimbJob J = desc.job(ct.y);
  const double q = P.qd[J.prime], qr = P.qr[J.prime];
  const ModF m{q, qr};
  const uint32_t *__restrict__ w = tw + (size_t)J.prime * kN;
  const int cc = threadIdx.x & 15;
  const int rb = threadIdx.x >> 4;  // 0..15

  // phase 1 twiddles psi_rev[2^s + blk] (uniform), phase 2 psi_rev[2^(4+s)
  // + rb 2^s + blk]
  uint32_t t1[15], t2[15];
  auto load_twiddles = [&]() {
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int blk = 0; blk < (1 << s); blk++) {
        t1[(1 << s) - 1 + blk] = w[(1 << s) + blk];
        t2[(1 << s) - 1 + blk] = w[(16 << s) + (rb << s) + blk];
      }
  };
  // phase 1 holds rows rb + 16 k of column c0 + cc
  double x[16];
  if (J.flags & kNttBconvInput) {
    // twiddles after the BaseConv, whose loads need the registers
    bconv_cols_to_lds<WIDTH>(J.bsrc, J.bs, J.btab, c0, m, lds);
    load_twiddles();
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = (double)lds[(rb + 16 * k) * kColTile + cc];
    __syncthreads();  // the staging words are reused by the double tile
  } else {
    load_twiddles();
    // canonical residues (< 2^31) or a lazy signed BaseConv output: int32
    const int32_t *__restrict__ g =
        reinterpret_cast<const int32_t *>(J.src) + c0 + cc;
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = (double)g[(rb + 16 * k) * 256];
  }
  if (J.flags & kNttPreMul) {
    const double cv = (double)J.c1, cq = cv * qr;
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = mulmod_f(x[k], cv, cq, q);
  }
  ct16(x, t1, q, qr);
#pragma unroll
  for (int k = 0; k < 16; k++) ldsd[(rb + 16 * k) * kColTile + cc] = x[k];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; j++) x[j] = ldsd[(rb * 16 + j) * kColTile + cc];
  ct16(x, t2, q, qr);
  // rows rb * 16 + j, lazy int32 for the row pass
  uint32_t *__restrict__ o = J.dst + c0 + cc;
#pragma unroll
  for (int j = 0; j < 16; j++)
    o[(rb * 16 + j) * 256] = (uint32_t)lazy_i32(x[j], m);
}

// ---------------------------------------------------------------------------
// Forward, row stages s = 8..15 (+ optional ModDown epilogue)
// ---------------------------------------------------------------------------
template <class D>
__global__ __launch_bounds__(kRowThreads, 5) void ntt_fwd_rows(
    const D desc, const PrimeTab P, const uint32_t *__restrict__ tw) {
  __shared__ double ldsd[kRowTile * 256];
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
  double x[16];
  rows_load_p1(J.dst + tile, rl, t16, x);
  // the ModDown epilogue's second operand, issued with the data loads so its
  // latency overlaps the row stages. Unconditional, since a conditional form
  // costs a spill here: without an epilogue every lane reads the same first
  // 64 bytes of the limb (one cached line per wave), unused.
  const bool has2 = (J.flags & kNttModDownEpilogue) && J.src2;
  const uint4 *s2p = reinterpret_cast<const uint4 *>(
      has2 ? J.src2 + (size_t)r * 256 + t16 * 16 : J.dst);
  uint4 s2[4];
#pragma unroll
  for (int i = 0; i < 4; i++) s2[i] = s2p[i];
  ct16(x, t1, q, qr);
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
    if (has2) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const uint4 u = s2[i];
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
      out[j] = canon_small(mulmod_f(a2[j] - x[j], ip, ipq, q), q);
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
__global__ __launch_bounds__(kRowThreads, 5) void intt_rows(
    const D desc, const PrimeTab P, const uint32_t *__restrict__ tw) {
  __shared__ double ldsd[kRowTile * 256];
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
  // GS sums double every stage: from inputs below 2^31 they stay below 2^39
  // over all eight stages, well inside mulmod_f's range, so the hand-off
  // needs no reduction
#pragma unroll
  for (int j = 0; j < 16; j++) ldsd[rlds(rl, 16 * t16 + j)] = x[j];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = ldsd[rlds(rl, t16 + 16 * k)];
  gs16(x, t2, q, qr);
  int32_t o[16];
#pragma unroll
  for (int k = 0; k < 16; k++) o[k] = lazy_i32(x[k], m);
  rows_store_p1(J.dst + tile, rl, t16, o);
}

// ---------------------------------------------------------------------------
// Inverse, column stages u = 8..15, then the per-limb constant
// ---------------------------------------------------------------------------
template <class D>
__global__ __launch_bounds__(kColThreads, 5) void intt_cols(
    const D desc, const PrimeTab P, const uint32_t *__restrict__ tw) {
  __shared__ double ldsd[256 * kColTile];
  const ColTile ct = col_tile();  // same XCD-aware order as ntt_fwd_cols
  const int c0 = ct.c0;
  const LimbJob J = desc.job(ct.y);
  const double q = P.qd[J.prime], qr = P.qr[J.prime];
  const ModF m{q, qr};
  const uint32_t *__restrict__ w = tw + (size_t)J.prime * kN;
  const int cc = threadIdx.x & 15;
  const int rb = threadIdx.x >> 4;

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
  // the pass-1 intermediate (in dst) is lazy signed int32; phase 1 holds
  // rows rb * 16 + j of column c0 + cc
  double x[16];
  {
    const int32_t *__restrict__ g =
        reinterpret_cast<const int32_t *>(J.dst) + c0 + cc;
#pragma unroll
    for (int j = 0; j < 16; j++) x[j] = (double)g[(rb * 16 + j) * 256];
  }
  gs16(x, t1, q, qr);
  // lazy inputs (|x| < 0.75 q) grow below 2^39 over the eight stages
#pragma unroll
  for (int j = 0; j < 16; j++) ldsd[(rb * 16 + j) * kColTile + cc] = x[j];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = ldsd[(rb + 16 * k) * kColTile + cc];
  gs16(x, t2, q, qr);
  const double cv = (double)J.c1, cq = cv * qr;
  uint32_t *__restrict__ o = J.dst + c0 + cc;
  // x * c1 lands in (-0.75 q, 0.75 q): canonical [0, q) (above: never,
  // below 0: + q) or centred [-(q-1)/2, (q-1)/2] (one fix-up either side)
  // (integer fix-ups on the exact int32 conversion)
  const int32_t qi = (int32_t)P.q[J.prime];
  const bool centred = (J.flags & kInttCentred) != 0;
  const int32_t above = centred ? (qi - 1) >> 1 : qi - 1;
  const int32_t below = centred ? -((qi - 1) >> 1) : 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    int32_t v = (int32_t)mulmod_f(x[k], cv, cq, q);
    v = v > above ? v - qi : v;
    v = v < below ? v + qi : v;
    o[(rb + 16 * k) * 256] = (uint32_t)v;
  }
}

// ---------------------------------------------------------------------------
// Forward row stages fused with the key inner product (MultKeyNoModDown,
// src/core/Context.cpp:609-712: the ModUp's last NTT stages, then
// ElementWise::PAccum / CAccum, src/core/ElementWise.cu:225-366)
// ---------------------------------------------------------------------------
//
// One workgroup owns a 16-row tile of output limb y (prime p). For every
// ModUp group g it takes limb y of that group's ModUp output straight from
// the column pass (lazy int32 in `modup`), runs the row stages in registers
// and multiplies the 16 values each thread holds by the key limbs, so the
// ModUp output never goes back to memory in NTT form. A group's own limbs
// are the ModUp source itself (x, or the Tensor product a1 (*) a2). In the
// Montgomery convention of the stored data the result is
//   acc_x[y] = R^-1 (sum_g m_g * key_x,g + fold_x)  mod q,
// fold_b = P * b1 b2 (HMult) or P R * add_b, fold_a likewise, which is
// exactly the key_accum kernel's sum of Montgomery products; here every
// product is an exact FP64 modular product of lazy values and one product by
// R^-1 ends the sum.
template <class Args>
__device__ __forceinline__ void key_rows_body(const EwNp &np, const PrimeTab &P,
                                              Args &A, const uint32_t *tw, int y,
                                              double *ldsd, uint32_t *tws) {
  const int p = y < np.num_q ? np.q_prime0 + y : np.aux_prime0 + (y - np.num_q);
  const double q = P.qd[p], qr = P.qr[p];
  const uint32_t qu = P.q[p];
  const int32_t qinv = P.qinv[p];
  const uint32_t *__restrict__ w = tw + (size_t)p * kN;
  const int rl = threadIdx.x >> 4;
  const int t16 = threadIdx.x & 15;
  const int r = blockIdx.x * kRowTile + rl;
  const size_t tile = (size_t)blockIdx.x * kRowTile * 256;
  const size_t tot_stride = (size_t)np.num_total * kN;
  const size_t off = (size_t)r * 256 + t16 * 16;  // this thread's 16 outputs
  const size_t my = (size_t)y * kN + off;
  const int kl = y < np.num_q ? A.key_q0 + y : A.key_aux0 + (y - np.num_q);
  const size_t kp = (size_t)kl * kN + off;
  const bool tens = A.use_tensor && y < np.num_q;
  // R^-1 and P = (P R) R^-1 as plain residues (Montgomery reductions)
  const double rinv = (double)mont_reduce(1, qu, qinv), rinvq = rinv * qr;

  // row-stage twiddles (as ntt_fwd_rows), parked in LDS for the group loop:
  // slot i of this thread at tws[i * 256 + tid] (phase 1: 0..14, 2: 15..29);
  // only this thread reads its slots back
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int blk = 0; blk < (1 << s); blk++) {
      tws[((1 << s) - 1 + blk) * kRowThreads + threadIdx.x] = w[(256 << s) + (r << s) + blk];
      tws[(15 + (1 << s) - 1 + blk) * kRowThreads + threadIdx.x] =
          w[(4096 << s) + (r << (4 + s)) + (t16 << s) + blk];
    }

  double sb[16], sa[16];
#pragma unroll
  for (int j = 0; j < 16; j++) sb[j] = sa[j] = 0.0;
  // the column-pass tile of the next group that needs the row stages is
  // loaded one group ahead (into `pre`), so its latency overlaps the
  // current group's arithmetic
  auto next_rows_group = [&](int from) {
    for (int i = from; i < A.ngroups; i++) {
      const int g = A.group[i];
      if (!(y >= A.own0[g] && y < A.own1[g])) return i;
    }
    return -1;
  };
  typedef const __attribute__((address_space(1))) int32_t gint;
  int32_t pre[16];
  auto prefetch = [&](int gi) {
    gint *src = (gint *)(A.modup + (size_t)A.group[gi] * tot_stride + (size_t)y * kN + tile +
                         rl * 256 + t16);
#pragma unroll
    for (int k = 0; k < 16; k++) pre[k] = src[16 * k];
  };
  int pg = next_rows_group(0);
  if (pg >= 0) prefetch(pg);
  for (int gi = 0; gi < A.ngroups; gi++) {
    const int g = A.group[gi];
    const bool own = (y >= A.own0[g] && y < A.own1[g]);
    // key limbs first: their loads overlap the row stages
    uint4 kb[4], ka[4];
    const uint32_t *kbp = A.key_b[g] + kp, *kap = A.key_a[g] + kp;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      kb[i] = reinterpret_cast<const uint4 *>(kbp)[i];
      ka[i] = reinterpret_cast<const uint4 *>(kap)[i];
    }
    double x[16];
    if (own) {
      if (tens) {
        // d2 = a1 (*) a2 = a1 a2 R^-1
        const uint4 *a1 = reinterpret_cast<const uint4 *>(A.tensor[1] + my);
        const uint4 *a2 = reinterpret_cast<const uint4 *>(A.tensor[3] + my);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const uint4 u = a1[i], v = a2[i];
          const uint32_t U[4] = {u.x, u.y, u.z, u.w}, V[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
          for (int e = 0; e < 4; e++) {
            const double vd = (double)V[e];
            const double t = mulmod_fq((double)U[e], vd, q, qr);
            x[4 * i + e] = mulmod_f(t, rinv, rinvq, q);
          }
        }
      } else {
        const uint4 *xs = reinterpret_cast<const uint4 *>(A.x + my);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const uint4 u = xs[i];
          x[4 * i + 0] = (double)u.x;
          x[4 * i + 1] = (double)u.y;
          x[4 * i + 2] = (double)u.z;
          x[4 * i + 3] = (double)u.w;
        }
      }
    } else {
      // gi == pg: its tile is in `pre`; start the next one
#pragma unroll
      for (int k = 0; k < 16; k++) x[k] = (double)pre[k];
      pg = next_rows_group(gi + 1);
      if (pg >= 0) prefetch(pg);
      ct16_lds(x, tws, q, qr);
      __syncthreads();  // the previous group's exchange may still be read
#pragma unroll
      for (int k = 0; k < 16; k++) ldsd[rlds(rl, t16 + 16 * k)] = x[k];
      __syncthreads();
#pragma unroll
      for (int j = 0; j < 16; j++) x[j] = ldsd[rlds(rl, 16 * t16 + j)];
      ct16_lds(x, tws + 15 * kRowThreads, q, qr);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const uint32_t KB[4] = {kb[i].x, kb[i].y, kb[i].z, kb[i].w};
      const uint32_t KA[4] = {ka[i].x, ka[i].y, ka[i].z, ka[i].w};
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const double b = (double)KB[e], a = (double)KA[e];
        sb[4 * i + e] += mulmod_fq(x[4 * i + e], b, q, qr);
        sa[4 * i + e] += mulmod_fq(x[4 * i + e], a, q, qr);
      }
    }
  }

  // the P fold on q limbs
  if (tens) {
    const double pq = (double)mont_reduce(A.p_prod[p], qu, qinv), pqq = pq * qr;
    const uint4 *b1 = reinterpret_cast<const uint4 *>(A.tensor[0] + my);
    const uint4 *a1 = reinterpret_cast<const uint4 *>(A.tensor[1] + my);
    const uint4 *b2 = reinterpret_cast<const uint4 *>(A.tensor[2] + my);
    const uint4 *a2 = reinterpret_cast<const uint4 *>(A.tensor[3] + my);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const uint4 u1 = b1[i], v1 = a1[i], u2 = b2[i], v2 = a2[i];
      const uint32_t B1[4] = {u1.x, u1.y, u1.z, u1.w}, A1[4] = {v1.x, v1.y, v1.z, v1.w};
      const uint32_t B2[4] = {u2.x, u2.y, u2.z, u2.w}, A2[4] = {v2.x, v2.y, v2.z, v2.w};
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const double db1 = (double)B1[e], da1 = (double)A1[e];
        const double db2 = (double)B2[e], da2 = (double)A2[e];
        const double d0 = mulmod_fq(db1, db2, q, qr);
        const double d1 = mulmod_fq(db1, da2, q, qr) + mulmod_fq(da1, db2, q, qr);
        sb[4 * i + e] += mulmod_f(d0, pq, pqq, q);
        sa[4 * i + e] += mulmod_f(d1, pq, pqq, q);
      }
    }
  } else if (y < np.num_q && (A.add_b || A.add_a)) {
    const double pp = (double)A.p_prod[p], ppq = pp * qr;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      if (A.add_b) {
        const uint4 u = reinterpret_cast<const uint4 *>(A.add_b + my)[i];
        const uint32_t U[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int e = 0; e < 4; e++) sb[4 * i + e] += mulmod_f((double)U[e], pp, ppq, q);
      }
      if (A.add_a) {
        const uint4 u = reinterpret_cast<const uint4 *>(A.add_a + my)[i];
        const uint32_t U[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int e = 0; e < 4; e++) sa[4 * i + e] += mulmod_f((double)U[e], pp, ppq, q);
      }
    }
  }

  uint4 *ob = reinterpret_cast<uint4 *>(A.out_b + my);
  uint4 *oa = reinterpret_cast<uint4 *>(A.out_a + my);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    uint32_t vb[4], va[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
      vb[e] = canon_small(mulmod_f(sb[4 * i + e], rinv, rinvq, q), q);
      va[e] = canon_small(mulmod_f(sa[4 * i + e], rinv, rinvq, q), q);
    }
    ob[i] = make_uint4(vb[0], vb[1], vb[2], vb[3]);
    oa[i] = make_uint4(va[0], va[1], va[2], va[3]);
  }
}

__global__ __launch_bounds__(kRowThreads, 2) void key_rows_kernel(
    const EwNp np, const PrimeTab P, const KeyAccumArgs A, const uint32_t *__restrict__ tw) {
  __shared__ double ldsd[kRowTile * 256];
  __shared__ uint32_t tws[30 * kRowThreads];
  key_rows_body(np, P, A, tw, blockIdx.y, ldsd, tws);
}

__global__ __launch_bounds__(kRowThreads, 2) void key_rows_table_kernel(
    const KeyAccumItem *__restrict__ items, const EwJob *__restrict__ jobs,
    const PrimeTab P, const uint32_t *__restrict__ tw) {
  __shared__ double ldsd[kRowTile * 256];
  __shared__ uint32_t tws[30 * kRowThreads];
  typedef const __attribute__((address_space(4))) KeyAccumItem CItem;
  typedef const __attribute__((address_space(4))) EwJob CJob;
  CJob &jb = ((CJob *)jobs)[blockIdx.y];
  const int item = jb.item, y = jb.y;
  CItem &it = ((CItem *)items)[item];
  const EwNp np{it.np.num_q, it.np.num_total, it.np.q_prime0, it.np.aux_prime0};
  key_rows_body(np, P, it.A, tw, y, ldsd, tws);
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
void fwd_passes(const Context &ctx, const D &d, int n, int flags, int max_bs,
                hipStream_t st) {
  if ((flags & kNttBconvInput) && max_bs > 16)
    launch(ctx, fwd_cols_name(flags), n, ntt_fwd_cols<D, 2>, dim3(kColTiles * n),
           dim3(kColThreads), st, d, ctx.ptab, (const uint32_t *)ctx.d_tw_fwd);
  else if ((flags & kNttBconvInput) && max_bs > 9)
    launch(ctx, fwd_cols_name(flags), n, ntt_fwd_cols<D, 1>, dim3(kColTiles * n),
           dim3(kColThreads), st, d, ctx.ptab, (const uint32_t *)ctx.d_tw_fwd);
  else
    launch(ctx, fwd_cols_name(flags), n, ntt_fwd_cols<D, 0>, dim3(kColTiles * n),
           dim3(kColThreads), st, d, ctx.ptab, (const uint32_t *)ctx.d_tw_fwd);
  if (!(flags & kNttColsOnly))
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
  int max_bs = 0;
  for (int i = 0; i < 4; i++) max_bs = std::max(max_bs, L.bs[i]);
  fwd_passes(ctx, JobByValue{L}, L.n, L.flags, max_bs, st);
}

// Inverse NTT of L.n limbs: row stages into dst, then column stages in
// place on dst with the post constant c1 (plain form) and optional centring.
void launch_intt(const Context &ctx, const NttLaunch &L, hipStream_t st) {
  if (L.n <= 0) return;
  inv_passes(ctx, JobByValue{L}, L.n, L.flags, st);
}

void launch_ntt_fwd_table(const Context &ctx, const LimbJob *jobs, int n,
                          int flags_any, int max_bs, hipStream_t st) {
  if (n <= 0) return;
  fwd_passes(ctx, JobTable{jobs}, n, flags_any, max_bs, st);
}

void launch_key_rows(const Context &ctx, const EwNp &np, const KeyAccumArgs &A,
                     hipStream_t st) {
  if (np.num_total <= 0) return;
  launch(ctx, "ntt_fwd_rows+key", np.num_total, key_rows_kernel, dim3(kRowTiles, np.num_total),
         dim3(kRowThreads), st, np, ctx.ptab, A, (const uint32_t *)ctx.d_tw_fwd);
}

void launch_key_rows_table(const Context &ctx, const KeyAccumItem *items, const EwJob *jobs,
                           int njobs, hipStream_t st) {
  if (njobs <= 0) return;
  launch(ctx, "ntt_fwd_rows+key[batch]", njobs, key_rows_table_kernel, dim3(kRowTiles, njobs),
         dim3(kRowThreads), st, items, jobs, ctx.ptab, (const uint32_t *)ctx.d_tw_fwd);
}

void launch_intt_table(const Context &ctx, const LimbJob *jobs, int n,
                       int flags_any, hipStream_t st) {
  if (n <= 0) return;
  inv_passes(ctx, JobTable{jobs}, n, flags_any, st);
}

}  // namespace ck
