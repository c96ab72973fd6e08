// Exact int8 matrix-core forms of the forward negacyclic NTT (4-step,
// 256 x 256) and of the fast BaseConv, plus a vector-unit BaseConv.
//
// Byte splitting. A residue x with |x| < 2^31 - 2^24 has a unique
// balanced base-256 form x = sum_i d_i 2^(8 i), d_i in [-128, 127]: the
// bytes of x + 0x80808080 with their top bit flipped (two instructions,
// pack_c). For a constant matrix T over Z_q the host stores, for every data
// digit i, T^(i) = T 2^(8 i) mod q (centred) split the same way into four
// byte planes T^(i)_j. Then
//   sum_n T[m][n] x_n  ==  sum_j 2^(8 j) D_j[m]   (mod q),
//   D_j[m] = sum_{n, i} T^(i)_j[m][n] d_i(x_n),
// and each D_j is one int8 matrix product with int32 accumulation over
// K = (input, digit). With K = 64 every |D_j| <= 64 * 2^14 = 2^20, so
// lo = D_0 + 2^8 D_1 and hi = D_2 + 2^8 D_3 stay within int32 and
// v = hi 2^16 + lo (|v| < 2^45) is exact in FP64; one FP64 reduction
// (k = rint(v / q), r = v - k q, exact) gives |r| <= q/2 + 1. No step
// rounds: the outputs are the exact residues.
//
// Operand layouts (gfx950, checked on the GPU): for
// v_mfma_i32_16x16x64_i8, lane l supplies A[l % 16][16 (l / 16) + b] and
// B[16 (l / 16) + b][l % 16] for byte b of its four registers and receives
// D[4 (l / 16) + v][l % 16] in accumulator v; v_mfma_i32_16x16x32_i8 is the
// same with eight bytes per lane. A lane's register t holds the four digits
// of one value (input 4 (l / 16) + t), so the reduction index k = (input,
// digit) pairs the same digit positions in A and B.
//
// 4-step transform. A limb x is the 256 x 256 matrix X[r][c] = x[c + 256 r].
// With psi the context's 2N-th root, psi1 = psi^256, omega = psi^512 and
// output position p = p_lo + 256 p_hi holding x(psi^(2 brv16(p) + 1)) (the
// order of NTTPhase1/NTTPhase2, src/core/NTT.cu:231-454):
//   column pass  Y[p_hi][c] = sum_r psi1^(r (2 brv8(p_hi) + 1)) X[r][c],
//                 Z = Y psi^(c (2 brv8(p_hi) + 1))        (4-step twiddle),
//   row pass     out[p_hi][p_lo] = sum_c omega^(c brv8(p_lo)) Z[p_hi][c].
// Each 256-point matrix is applied as two 16-point matrix stages
// (r = r0 + 16 r1, p_hi = p0 + 16 p1; c = c0 + 16 c1, p_lo = q0 + 16 q1):
//   A1[p1][r1]     = psi1^(16 r1 (2 brv4(p1) + 1)),
//   B1_p1[p0][r0]  = psi1^(r0 (2 brv4(p1) + 1) + 32 r0 brv4(p0)),
//   A3[q1][c1]     = omega^(16 c1 brv4(q1)),
//   B3_q1[q0][c0]  = omega^(c0 brv4(q1) + 16 c0 brv4(q0)),
// the inner twiddles folded into the second stage's 16 matrices. Between
// the two stages of a pass the values go through LDS (the matrix cores
// leave a stage's output index where the next stage needs its input index).
// The column pass reads canonical input as x - 2^30 (fits the digit range)
// and adds 2^30 * rowsum(A1) back through the accumulator's initial value.
// Intermediate values between the passes are centred int32 in dst, in
// 16 x 16 tiles (tile_word).
#include "ck_i8.h"
#include "ck_i8_arith.h"

#include <algorithm>
#include <cstring>

namespace ck {
namespace {

constexpr int kN = 65536;
constexpr int kXs = 324;  // LDS words per 16 x 16 block of the exchange

typedef int v4i __attribute__((ext_vector_type(4)));

using i8::combine;
using i8::pack_c;
using i8::pack_off;
using i8::reduce_q;

// workgroup b runs on XCD b % 8: give each XCD a contiguous range of
// virtual blocks, so workgroups of one prime (adjacent jobs) share an L2
__device__ __forceinline__ int xcd_block(int b, int nwg) {
  const int x = b & 7, i = b >> 3;
  const int per = nwg >> 3, rem = nwg & 7;
  return x * per + (x < rem ? x : rem) + i;
}

__device__ __forceinline__ v4i mfma64(v4i a, v4i b, v4i c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

typedef const __attribute__((address_space(4))) Ntt8Job CJob;
typedef uint32_t v4u __attribute__((ext_vector_type(4)));

// data that streams through once (the intermediate tiles, the outputs):
// non-temporal accesses keep it from evicting the twiddle table and operand
// matrices, which every limb of a prime re-reads. The column pass reads its
// input with ordinary loads: it uses each 128-byte line in two 64-byte
// halves, and a line kept in L2 serves the second half.
__device__ __forceinline__ v4u ld4_nt(const uint32_t *p) {
  return __builtin_nontemporal_load((const v4u *)p);
}
__device__ __forceinline__ void st_nt(uint32_t *p, uint32_t v) {
  __builtin_nontemporal_store(v, p);
}
__device__ __forceinline__ void st4_nt(uint32_t *p, v4u v) {
  __builtin_nontemporal_store(v, (v4u *)p);
}

// Intermediate layout between the passes: 16 x 16 tiles of 1 KiB, tile
// (p1, cb) = rows 16 p1 .. 16 p1 + 15, columns 16 cb .. 16 cb + 15 at word
// (16 p1 + cb) * 256; inside a tile, row p0 = 4 g + v and column n sit at
// word (16 g + n) * 4 + v. The column pass writes a tile with one 16-byte
// store per lane (1 KiB per wave), the row pass of rows 16 p1.. reads the
// 16 KiB of its 16 tiles contiguously.
__device__ __forceinline__ int tile_word(int p1, int cb, int g, int n) {
  return (16 * p1 + cb) * 256 + (16 * g + n) * 4;
}

// Four planes of one 16 x 16 output tile (matrix a[j] times packed data b,
// plane 0 starting from c0).
struct Acc4 {
  v4i d[4];
};
__device__ __forceinline__ Acc4 mfma_tile(const v4i *a, v4i b, v4i c0) {
  const v4i zero = {0, 0, 0, 0};
  Acc4 r;
  r.d[0] = mfma64(a[0], b, c0);
  r.d[1] = mfma64(a[1], b, zero);
  r.d[2] = mfma64(a[2], b, zero);
  r.d[3] = mfma64(a[3], b, zero);
  return r;
}
// data as the A operand (rows = data vectors), matrix as B
__device__ __forceinline__ Acc4 mfma_tile_t(v4i a, const v4i *b) {
  const v4i zero = {0, 0, 0, 0};
  Acc4 r;
  r.d[0] = mfma64(a, b[0], zero);
  r.d[1] = mfma64(a, b[1], zero);
  r.d[2] = mfma64(a, b[2], zero);
  r.d[3] = mfma64(a, b[3], zero);
  return r;
}
// the four outputs of a lane, recombined and reduced, each step over all
// four before the next (four independent dependency chains)
__device__ __forceinline__ void reduce4(const Acc4 &A, double q, double qr, double r[4]) {
  double v[4];
#pragma unroll
  for (int i = 0; i < 4; i++) v[i] = combine(A.d[0][i], A.d[1][i], A.d[2][i], A.d[3][i]);
#pragma unroll
  for (int i = 0; i < 4; i++) r[i] = reduce_q(v[i], q, qr);
}

// Column pass: workgroup part of a limb's `parts` = 16 / cbs workgroups
// takes the 16-column blocks part, part + parts, ...: with two parts the
// workgroups of a limb (adjacent, so on one XCD and started together) read
// the two 64-byte halves of each 128-byte input line at about the same
// time. Wave w runs the first stage for r0 = 4w..4w+3 and the second stage for
// p1 = 4w..4w+3, so it keeps A1 and four B1 matrices in registers. Each
// stage is software-pipelined over its four tiles: the next tile's matrix
// products (and LDS reads) are issued before the current tile's
// recombination, and the next block's inputs load during the block.
__global__ __launch_bounds__(256, 3) void ntt8_cols_kernel(const Ntt8Job *__restrict__ jobs,
                                                           int nwg, int cbs, I8NttTabs T,
                                                           PrimeTab P) {
  __shared__ __align__(16) int xs[2][16 * kXs];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int g = lane >> 4, n = lane & 15;
  const int vb = xcd_block(blockIdx.x, nwg);
  const int parts = 16 / cbs;
  const int job = vb / parts, part = vb - job * parts;
  CJob &J = ((CJob *)jobs)[job];
  const int pr = J.prime, mont = J.mont;
  const uint32_t *__restrict__ src = J.src;
  uint32_t *__restrict__ dst = J.dst;

  const v4i *pl = (const v4i *)T.planes + (size_t)pr * kI8Ctx * 256;
  const int ca = mont ? kI8CtxA1Mont : kI8CtxA1;
  v4i a1[4], b1[4][4];
#pragma unroll
  for (int j = 0; j < 4; j++) a1[j] = pl[(ca * 4 + j) * 64 + lane];
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int j = 0; j < 4; j++) b1[k][j] = pl[((kI8CtxB1 + 4 * w + k) * 4 + j) * 64 + lane];
  const int *cp = T.corr + (pr * 2 + mont) * 16 + 4 * g;
  const v4i corr = {cp[0], cp[1], cp[2], cp[3]};
  const v4i zero = {0, 0, 0, 0};
  const double q = P.qd[pr], qr = P.qr[pr];
  // the twiddle table is tiled like the intermediate: one 16-byte load
  // gives a lane the four twiddles of its outputs in tile (p1, cb)
  const uint32_t *__restrict__ tw0 = T.tw + (size_t)pr * kN + (16 * g + n) * 4;
  // lane's first input word of a block: row 4w + 64 g, column n
  const uint32_t *in0 = src + (4 * w + 64 * g) * 256 + n;
  // this lane's 16-byte slot of tile (p1, cb)
  uint32_t *out0 = dst + (16 * g + n) * 4;

  uint32_t raw[4][4];
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int t = 0; t < 4; t++) raw[k][t] = in0[(k + 16 * t) * 256 + part * 16];

  for (int it = 0; it < cbs; it++) {
    const int cb = part + parts * it;
    int *X = xs[it & 1];
    // first stage: 16-point transforms along r1, tile r0 = 4w + k
    Acc4 acc[2];
    {
      const v4i b = {pack_off(raw[0][0]), pack_off(raw[0][1]), pack_off(raw[0][2]),
                     pack_off(raw[0][3])};
      acc[0] = mfma_tile(a1, b, corr);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (k < 3) {
        const v4i b = {pack_off(raw[k + 1][0]), pack_off(raw[k + 1][1]),
                       pack_off(raw[k + 1][2]), pack_off(raw[k + 1][3])};
        acc[(k + 1) & 1] = mfma_tile(a1, b, corr);
      }
      double r[4];
      reduce4(acc[k & 1], q, qr, r);
#pragma unroll
      for (int v = 0; v < 4; v++) X[(4 * g + v) * kXs + (4 * w + k) * 20 + n] = pack_c((int)r[v]);
    }
    // next block's inputs, in flight during the second stage (the last
    // iteration reloads its own block)
    {
      const int cbn = it + 1 < cbs ? cb + parts : cb;
#pragma unroll
      for (int k = 0; k < 4; k++)
#pragma unroll
        for (int t = 0; t < 4; t++) raw[k][t] = in0[(k + 16 * t) * 256 + cbn * 16];
    }
    __syncthreads();
    // second stage along r0 (tile p1 = 4w + k), then the 4-step twiddle;
    // one 16-byte store per lane writes the tile (p1, cb)
    v4u twv[2];
    twv[0] = *(const v4u *)(tw0 + (16 * (4 * w) + cb) * 256);
    v4i bx[2];
    const int *Xl = X + (4 * w) * kXs + (4 * g) * 20 + n;
    bx[0] = v4i{Xl[0], Xl[20], Xl[40], Xl[60]};
    acc[0] = mfma_tile(b1[0], bx[0], zero);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (k < 3) {
        const int *Xk = Xl + (k + 1) * kXs;
        bx[(k + 1) & 1] = v4i{Xk[0], Xk[20], Xk[40], Xk[60]};
        twv[(k + 1) & 1] = *(const v4u *)(tw0 + (16 * (4 * w + k + 1) + cb) * 256);
        acc[(k + 1) & 1] = mfma_tile(b1[k + 1], bx[(k + 1) & 1], zero);
      }
      double r[4];
      reduce4(acc[k & 1], q, qr, r);
      v4u out;
#pragma unroll
      for (int v = 0; v < 4; v++)
        out[v] = (uint32_t)(int)mulmod_fq(r[v], (double)twv[k & 1][v], q, qr);
      st4_nt(out0 + (16 * (4 * w + k) + cb) * 256, out);
    }
  }
}

// Row pass (in place on dst): workgroup = one limb's 16-row blocks
// [part * rbs, +rbs). Wave w runs the first stage for rows 16 rb + 4w + k
// (one 16-byte load per lane and tile gives its four rows) and the second
// stage for q1 = 4w + k with the data as the A operand, so each lane ends
// with four rows of one output column. Software-pipelined like the column
// pass; the next block's tiles load while the current one is transformed.
__global__ __launch_bounds__(256, 3) void ntt8_rows_kernel(const Ntt8Job *__restrict__ jobs,
                                                           int nwg, int rbs, I8NttTabs T,
                                                           PrimeTab P) {
  __shared__ __align__(16) int xs[2][16 * kXs];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int g = lane >> 4, n = lane & 15;
  const int vb = xcd_block(blockIdx.x, nwg);
  const int parts = 16 / rbs;
  const int job = vb / parts, part = vb - job * parts;
  CJob &J = ((CJob *)jobs)[job];
  const int pr = J.prime;
  uint32_t *__restrict__ Z = J.dst;

  const v4i *pl = (const v4i *)T.planes + (size_t)pr * kI8Ctx * 256;
  v4i a3[4], b3[4][4];
#pragma unroll
  for (int j = 0; j < 4; j++) a3[j] = pl[(kI8CtxA3 * 4 + j) * 64 + lane];
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int j = 0; j < 4; j++) b3[k][j] = pl[((kI8CtxB3 + 4 * w + k) * 4 + j) * 64 + lane];
  const v4i zero = {0, 0, 0, 0};
  const double q = P.qd[pr], qr = P.qr[pr];

  const int rb0 = part * rbs, rb1 = rb0 + rbs;
  // tile (rb, c1 = 4g + t), rows 4w + k (k = 0..3 in one load), column c0 = n
  const uint32_t *in0 = Z + tile_word(0, 4 * g, w, n);
  v4u raw[4];
#pragma unroll
  for (int t = 0; t < 4; t++) raw[t] = ld4_nt(in0 + (16 * rb0 + t) * 256);

  for (int rb = rb0; rb < rb1; rb++) {
    int *W = xs[rb & 1];
    // first stage along c1 (row 16 rb + 4w + k): V[q1 = 4g + v][c0 = n]
    Acc4 acc[2];
    {
      const v4i b = {pack_c((int)raw[0][0]), pack_c((int)raw[1][0]), pack_c((int)raw[2][0]),
                     pack_c((int)raw[3][0])};
      acc[0] = mfma_tile(a3, b, zero);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (k < 3) {
        const v4i b = {pack_c((int)raw[0][k + 1]), pack_c((int)raw[1][k + 1]),
                       pack_c((int)raw[2][k + 1]), pack_c((int)raw[3][k + 1])};
        acc[(k + 1) & 1] = mfma_tile(a3, b, zero);
      }
      double r[4];
      reduce4(acc[k & 1], q, qr, r);
#pragma unroll
      for (int v = 0; v < 4; v++) W[(4 * w + k) * kXs + (4 * g + v) * 20 + n] = pack_c((int)r[v]);
    }
    // next block's tiles, in flight during the second stage
    {
      const int rbn = rb + 1 < rb1 ? rb + 1 : rb;
#pragma unroll
      for (int t = 0; t < 4; t++) raw[t] = ld4_nt(in0 + (16 * rbn + t) * 256);
    }
    __syncthreads();
    // the row block's tiles are all read: the outputs may overwrite them
    // second stage along c0 (q1 = 4w + k): rows as A (row n, c0 = 4g + t)
    const v4i *Wl = (const v4i *)&W[n * kXs + (4 * w) * 20 + 4 * g];
    uint32_t *out0 = Z + (16 * rb + 4 * g) * 256 + n + 16 * (4 * w);
    v4i ax[2];
    ax[0] = Wl[0];
    acc[0] = mfma_tile_t(ax[0], b3[0]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (k < 3) {
        ax[(k + 1) & 1] = Wl[5 * (k + 1)];  // + (k + 1) * 20 words
        acc[(k + 1) & 1] = mfma_tile_t(ax[(k + 1) & 1], b3[k + 1]);
      }
      double r[4];
      reduce4(acc[k & 1], q, qr, r);
#pragma unroll
      for (int v = 0; v < 4; v++) st_nt(out0 + v * 256 + 16 * k, canon_small(r[v], q));
    }
  }
}

// ---- BaseConv ---------------------------------------------------------------

typedef const __attribute__((address_space(4))) BconvI8Inst CBInst;

// int8 form: workgroup = 256 positions of one instance (64 per wave, four
// 16-position tiles). Rows = 16 destinations per tile, K = 8 sources x 4
// digits per v_mfma_i32_16x16x32_i8 (KS <= 3 k-steps for up to 8 KS
// sources: each plane sums at most 96 digit products, |D_j| < 2^21).
// Source indices past S are clamped to 0 (their plane bytes are 0), so every
// load is unconditional.
template <int KS, class Inst>
__device__ __forceinline__ void bconv_i8_body(Inst &I, const PrimeTab &P) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int g = lane >> 4, n = lane & 15;
  const int S = I.S, D = I.D, canonical = I.canonical;
  const int base = blockIdx.x * 256 + w * 64 + n;
  const int32_t *__restrict__ src = I.src;
  const int64_t *__restrict__ planes = I.planes;

  int raw[4][KS][2];
#pragma unroll
  for (int ks = 0; ks < KS; ks++)
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const int k = 8 * ks + 2 * g + e;
      const int32_t *sp = src + (size_t)(k < S ? k : 0) * kN + base;
#pragma unroll
      for (int t = 0; t < 4; t++) raw[t][ks][e] = sp[16 * t];
    }
  int64_t b[4][KS];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int ks = 0; ks < KS; ks++)
      b[t][ks] = (int64_t)(uint32_t)pack_c(raw[t][ks][0]) |
                 ((int64_t)(uint32_t)pack_c(raw[t][ks][1]) << 32);

  const int nks = (S + 7) >> 3;  // this instance's k-steps (plane layout)
  const v4i zero = {0, 0, 0, 0};
  for (int mt = 0; mt * 16 < D; mt++) {
    int64_t a[KS][4];
#pragma unroll
    for (int ks = 0; ks < KS; ks++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        a[ks][j] = ks < nks ? planes[((mt * nks + ks) * 4 + j) * 64 + lane] : 0;
    double qv[4], qrv[4];
    uint32_t *dp[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int d = mt * 16 + 4 * g + v;
      const bool ok = d < D;
      const int dd = ok ? d : 0;
      const int p = I.dst_prime[dd];
      qv[v] = P.qd[p];
      qrv[v] = P.qr[p];
      dp[v] = ok ? I.dst + (size_t)I.dst_limb[dd] * kN : nullptr;
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
      v4i d[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        v4i acc = zero;
#pragma unroll
        for (int ks = 0; ks < KS; ks++)
          acc = __builtin_amdgcn_mfma_i32_16x16x32_i8(a[ks][j], b[t][ks], acc, 0, 0, 0);
        d[j] = acc;
      }
#pragma unroll
      for (int v = 0; v < 4; v++) {
        if (dp[v]) {
          const double r = reduce_q(combine(d[0][v], d[1][v], d[2][v], d[3][v]), qv[v], qrv[v]);
          dp[v][base + 16 * t] = canonical ? canon_small(r, qv[v]) : (uint32_t)(int)r;
        }
      }
    }
  }
}

template <int KS>
__global__ __launch_bounds__(256) void bconv_i8_kernel(const BconvI8Inst *__restrict__ insts,
                                                       PrimeTab P) {
  CBInst &I = ((CBInst *)insts)[blockIdx.y];
  bconv_i8_body<KS>(I, P);
}

template <int KS>
__global__ __launch_bounds__(256) void bconv_i8_arg_kernel(const BconvI8Launch L,
                                                           PrimeTab P) {
  bconv_i8_body<KS>(L.inst[blockIdx.y], P);
}

// Vector-unit form: one position per thread, the table row of each
// destination from scalar loads; per destination sum_i v_i hi_i and
// sum_i v_i lo_i in exact FP64 (|v| < 2^30, halves < 2^16, S <= 24 terms:
// below 2^51), then hi 2^16 + lo reduced as in the fused column pass.
__global__ __launch_bounds__(256) void bconv_valu_kernel(const BconvI8Inst *__restrict__ insts,
                                                         PrimeTab P) {
  CBInst &I = ((CBInst *)insts)[blockIdx.y];
  const int pos = blockIdx.x * 256 + threadIdx.x;
  const int S = I.S, D = I.D, canonical = I.canonical;
  double v[kMaxBconvSrc];
#pragma unroll
  for (int i = 0; i < kMaxBconvSrc; i++) {
    const int32_t x = I.src[(size_t)(i < S ? i : 0) * kN + pos];
    v[i] = i < S ? (double)x : 0.0;
  }
  typedef const __attribute__((address_space(4))) double CD;
  for (int d = 0; d < D; d++) {
    CD *row = (CD *)I.ftab + (size_t)d * S * 2;
    double hi = 0.0, lo = 0.0;
#pragma unroll
    for (int i = 0; i < kMaxBconvSrc; i++) {
      if (i < S) {
        hi = __builtin_fma(v[i], row[2 * i], hi);
        lo = __builtin_fma(v[i], row[2 * i + 1], lo);
      }
    }
    const int p = I.dst_prime[d];
    const ModF mf{P.qd[p], P.qr[p]};
    const double big = hi * 65536.0;
    const double k = __builtin_rint(big * mf.qinv);
    const double r = reduce_f(__builtin_fma(-k, mf.q, big) + lo, mf);
    I.dst[(size_t)I.dst_limb[d] * kN + pos] =
        canonical ? canon_small(r, mf.q) : (uint32_t)(int32_t)r;
  }
}

// ---- host tables ------------------------------------------------------------

uint32_t brv_n(uint32_t x, int bits) {
  uint32_t r = 0;
  for (int i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}

}  // namespace

int ntt8_tables(Context &c, I8NttTabs *out) {
  if (c.i8_planes) {
    *out = I8NttTabs{c.i8_planes, c.i8_corr, c.i8_tw};
    return kOk;
  }
  const int np = c.num_primes;
  std::vector<int32_t> planes((size_t)np * kI8Ctx * 1024);
  std::vector<int32_t> corr((size_t)np * 32);
  std::vector<uint32_t> tw((size_t)np * kN);
  uint32_t T[16][16];
  for (int pi = 0; pi < np; pi++) {
    const uint32_t q = c.primes[pi];
    const uint32_t psi = c.psi[pi];
    const uint32_t psi1 = h_powmod(psi, 256, q), om = h_powmod(psi, 512, q);
    const uint32_t rmod = (uint32_t)((1ull << 32) % q);
    int32_t *pp = planes.data() + (size_t)pi * kI8Ctx * 1024;
    // A1 (plain and times R) with its offset corrections
    for (int mont = 0; mont < 2; mont++) {
      for (int p1 = 0; p1 < 16; p1++)
        for (int r1 = 0; r1 < 16; r1++) {
          uint32_t v = h_powmod(psi1, 16ull * r1 * (2 * brv_n(p1, 4) + 1), q);
          T[p1][r1] = mont ? h_mulmod(v, rmod, q) : v;
        }
      i8::ctx_planes(T, q, pp + (size_t)(kI8CtxA1 + mont) * 1024);
      for (int m = 0; m < 16; m++) {
        uint64_t s = 0;
        for (int r1 = 0; r1 < 16; r1++) s += T[m][r1];
        corr[(size_t)pi * 32 + mont * 16 + m] =
            i8::centre_h((uint64_t)h_mulmod((uint32_t)(s % q), (1u << 30) % q, q), q);
      }
    }
    for (int p1 = 0; p1 < 16; p1++) {
      for (int p0 = 0; p0 < 16; p0++)
        for (int r0 = 0; r0 < 16; r0++)
          T[p0][r0] = h_powmod(psi1, (uint64_t)r0 * (2 * brv_n(p1, 4) + 1) +
                                         32ull * r0 * brv_n(p0, 4), q);
      i8::ctx_planes(T, q, pp + (size_t)(kI8CtxB1 + p1) * 1024);
    }
    for (int q1 = 0; q1 < 16; q1++)
      for (int c1 = 0; c1 < 16; c1++) T[q1][c1] = h_powmod(om, 16ull * c1 * brv_n(q1, 4), q);
    i8::ctx_planes(T, q, pp + (size_t)kI8CtxA3 * 1024);
    for (int q1 = 0; q1 < 16; q1++) {
      for (int q0 = 0; q0 < 16; q0++)
        for (int c0 = 0; c0 < 16; c0++)
          T[q0][c0] = h_powmod(om, (uint64_t)c0 * brv_n(q1, 4) + 16ull * c0 * brv_n(q0, 4), q);
      i8::ctx_planes(T, q, pp + (size_t)(kI8CtxB3 + q1) * 1024);
    }
    // 4-step twiddle psi^(c (2 brv8(p) + 1)) of row p, column c, stored in
    // the tiled layout of the intermediate (tile_word)
    uint32_t *tp = tw.data() + (size_t)pi * kN;
    for (int p = 0; p < 256; p++) {
      const uint32_t step = h_powmod(psi, 2 * brv_n(p, 8) + 1, q);
      uint64_t x = 1;
      const int p1 = p >> 4, p0 = p & 15;
      for (int col = 0; col < 256; col++) {
        const int cb = col >> 4, cn = col & 15;
        tp[(16 * p1 + cb) * 256 + (16 * (p0 >> 2) + cn) * 4 + (p0 & 3)] = (uint32_t)x;
        x = x * step % q;
      }
    }
  }
  auto up = [&](auto **dst, const auto &v) {
    int rc = check_hip(hipMalloc((void **)dst, v.size() * sizeof(v[0])), "hipMalloc(int8 tables)");
    if (rc) return rc;
    c.allocations.push_back((void *)*dst);
    return check_hip(hipMemcpy((void *)*dst, v.data(), v.size() * sizeof(v[0]),
                               hipMemcpyHostToDevice),
                     "hipMemcpy(int8 tables)");
  };
  int32_t *dp = nullptr, *dc = nullptr;
  uint32_t *dt = nullptr;
  int rc = up(&dp, planes);
  if (!rc) rc = up(&dc, corr);
  if (!rc) rc = up(&dt, tw);
  if (rc) return rc;
  c.i8_planes = dp;
  c.i8_corr = dc;
  c.i8_tw = dt;
  *out = I8NttTabs{dp, dc, dt};
  return kOk;
}

void launch_ntt8_cols(const Context &c, const I8NttTabs &T, const Ntt8Job *jobs, int n,
                      int col_blocks, hipStream_t st) {
  if (n <= 0) return;
  const int nwg = n * (16 / col_blocks);
  launch(c, "ntt8_cols", n, ntt8_cols_kernel, dim3(nwg), dim3(256), st, jobs, nwg,
         col_blocks, T, c.ptab);
}

void launch_ntt8_rows(const Context &c, const I8NttTabs &T, const Ntt8Job *jobs, int n,
                      int row_blocks, hipStream_t st) {
  if (n <= 0) return;
  const int nwg = n * (16 / row_blocks);
  launch(c, "ntt8_rows", n, ntt8_rows_kernel, dim3(nwg), dim3(256), st, jobs, nwg,
         row_blocks, T, c.ptab);
}

void launch_bconv_i8_table(const Context &c, const BconvI8Inst *insts, int n, int dsts,
                           int max_s, int, hipStream_t st) {
  if (n <= 0) return;
  if (max_s <= 8)
    launch(c, "bconv_i8", dsts, bconv_i8_kernel<1>, dim3(kN / 256, n), dim3(256), st, insts,
           c.ptab);
  else if (max_s <= 16)
    launch(c, "bconv_i8", dsts, bconv_i8_kernel<2>, dim3(kN / 256, n), dim3(256), st, insts,
           c.ptab);
  else
    launch(c, "bconv_i8", dsts, bconv_i8_kernel<3>, dim3(kN / 256, n), dim3(256), st, insts,
           c.ptab);
}

void launch_bconv_i8(const Context &c, const BconvI8Launch &L, hipStream_t st) {
  if (L.n <= 0) return;
  int dsts = 0, max_s = 1;
  for (int i = 0; i < L.n; i++) {
    dsts += L.inst[i].D;
    max_s = std::max(max_s, L.inst[i].S);
  }
  if (max_s <= 8)
    launch(c, "bconv_i8", dsts, bconv_i8_arg_kernel<1>, dim3(kN / 256, L.n), dim3(256), st,
           L, c.ptab);
  else if (max_s <= 16)
    launch(c, "bconv_i8", dsts, bconv_i8_arg_kernel<2>, dim3(kN / 256, L.n), dim3(256), st,
           L, c.ptab);
  else
    launch(c, "bconv_i8", dsts, bconv_i8_arg_kernel<3>, dim3(kN / 256, L.n), dim3(256), st,
           L, c.ptab);
}

const int64_t *bconv_i8_planes_for(const Context &c, const double *d_ftab, int S, int D,
                                   const uint8_t *dst_prime) {
  for (auto &e : c.i8_bconv_planes)
    if (e.first == d_ftab) return e.second;
  std::vector<double> f((size_t)2 * S * D);
  if (check_hip(hipMemcpy(f.data(), d_ftab, f.size() * sizeof(double), hipMemcpyDeviceToHost),
                "hipMemcpy(BaseConv table)"))
    return nullptr;
  std::vector<uint32_t> tab((size_t)S * D), dq(D);
  for (int d = 0; d < D; d++) {
    dq[d] = c.primes[dst_prime[d]];
    for (int i = 0; i < S; i++)
      tab[(size_t)d * S + i] = (uint32_t)f[2 * ((size_t)d * S + i)] * 65536u +
                               (uint32_t)f[2 * ((size_t)d * S + i) + 1];
  }
  const std::vector<int64_t> pl = i8::bconv_i8_planes(tab, S, D, dq);
  int64_t *dp = nullptr;
  if (check_hip(hipMalloc((void **)&dp, pl.size() * sizeof(int64_t)), "hipMalloc(int8 planes)"))
    return nullptr;
  if (check_hip(hipMemcpy(dp, pl.data(), pl.size() * sizeof(int64_t), hipMemcpyHostToDevice),
                "hipMemcpy(int8 planes)")) {
    (void)hipFree(dp);
    return nullptr;
  }
  c.i8_bconv_planes.emplace_back(d_ftab, dp);
  return dp;
}

bool to_i8_inst(const Context &c, const BconvInst &in, BconvI8Inst *out) {
  std::memset(out, 0, sizeof(*out));
  out->planes = bconv_i8_planes_for(c, in.tab, in.S, in.D, in.dst_prime);
  if (!out->planes) return false;
  out->src = in.src;
  out->dst = in.dst;
  out->S = in.S;
  out->D = in.D;
  out->canonical = in.canonical;
  for (int m = 0; m < in.D; m++) {
    out->dst_limb[m] = in.dst_limb[m];
    out->dst_prime[m] = in.dst_prime[m];
  }
  return true;
}

void launch_bconv_valu_table(const Context &c, const BconvI8Inst *insts, int n, int dsts,
                             int, hipStream_t st) {
  if (n <= 0) return;
  launch(c, "bconv_valu", dsts, bconv_valu_kernel, dim3(kN / 256, n), dim3(256), st, insts,
         c.ptab);
}

}  // namespace ck
