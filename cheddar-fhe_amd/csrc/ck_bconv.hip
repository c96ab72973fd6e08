// BaseConv on the FP64 matrix cores (ModSwitchMatrixMult,
// src/core/ModSwitch.cu:23-171): for every coefficient position x and
// destination limb d,
//   y_d(x) = sum_k v_k(x) * T_dk  mod t_d,   T_dk = hi * 2^16 + lo,
// over S centred int32 source limbs v_k (|v| < 2^30). As a matrix product
// the rows are (destination, half) pairs, the reduction runs over the
// sources and the columns over positions; with 16-bit table halves every
// product is below 2^46 and every sum of S <= 16 products below 2^50, so
// v_mfma_f64_16x16x4_f64 computes the two partial sums exactly. The vector
// unit only recombines hi * 2^16 + lo and reduces (one exact modular
// reduction per output), instead of 2 S FMAs per output in the fused
// column pass.
//
// MFMA operand layout (16x16x4 f64, one wave): lane l supplies A[l%16][l/16]
// and B[l/16][l%16], and receives D[4i + l/16][l%16] in accumulator i. Row
// 4i + g of a 16-row tile is assigned to destination 2g + i/2 (of the tile's
// eight) and half i%2, so each lane ends with (hi, lo) of two destinations
// at one position and needs no cross-lane exchange.
//
// Output: lazy int32 in (-0.75 t, 0.75 t) per destination limb, the input
// the forward column pass reads.
#include "ck_bconv.h"

namespace ck {
namespace {

constexpr int kN = 65536;
constexpr int kThreads = 256;
constexpr int kChunk = 512;  // positions per workgroup (128 per wave)

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kTiles = kChunk / 4 / 16;  // 16-position tiles per wave

// I: the instance, read in place (kernel argument or constant-space table);
// RT / KS: static bounds on the row tiles (8 destinations each) and k-steps
// (4 sources each) so that the operand arrays stay in registers
template <int RT, int KS, class Inst>
__device__ __forceinline__ void bconv_body(Inst &I, const PrimeTab &P, int chunk) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4, n16 = lane & 15;
  const int S = I.S, D = I.D;
  const int nks = (S + 3) >> 2, nrt = (D + 7) >> 3;
  const int base = chunk * kChunk + wave * (kChunk / 4);

  // every source value this lane needs, loaded up front (one memory
  // latency per wave): source k = 4 ks + l / 16 at position base + 16 t + l % 16
  int32_t raw[kTiles][KS];
#pragma unroll
  for (int t = 0; t < kTiles; t++)
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      const int k = ks * 4 + g;
      raw[t][ks] = (ks < nks && k < S) ? I.src[(size_t)k * kN + base + t * 16 + n16] : 0;
    }

  // A operands of every (row tile, k-step): lane row m = l % 16 is
  // destination 2 (m % 4) + m / 8 of the tile, half (m / 4) % 2
  const int m = n16;
  const int dl = 2 * (m & 3) + (m >> 3), half = (m >> 2) & 1;
  double a[RT][KS];
#pragma unroll
  for (int rt = 0; rt < RT; rt++)
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      const int d = rt * 8 + dl, k = ks * 4 + g;
      a[rt][ks] = (rt < nrt && ks < nks && d < D && k < S)
                      ? I.tab[(d * S + k) * 2 + half]
                      : 0.0;
    }

  // per (row tile, e): this lane's destination 2g + e of the tile is the
  // same for every position tile, so its prime and limb are fetched once
  double eq[RT][2], eqr[RT][2];
  uint32_t *edst[RT][2];
#pragma unroll
  for (int rt = 0; rt < RT; rt++)
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const int d = rt * 8 + 2 * g + e;
      const bool ok = rt < nrt && d < D;
      const int p = ok ? I.dst_prime[d] : 0;
      eq[rt][e] = P.qd[p];
      eqr[rt][e] = P.qr[p];
      edst[rt][e] = ok ? I.dst + (size_t)I.dst_limb[d] * kN : nullptr;
    }

#pragma unroll
  for (int t = 0; t < kTiles; t++) {
    const int pos = base + t * 16 + n16;
    double b[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ks++) b[ks] = (double)raw[t][ks];
#pragma unroll
    for (int rt = 0; rt < RT; rt++) {
      if (rt >= nrt) break;
      v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        if (ks >= nks) break;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt][ks], b[ks], acc, 0, 0, 0);
      }
#pragma unroll
      for (int e = 0; e < 2; e++) {
        if (edst[rt][e]) {
          const ModF mf{eq[rt][e], eqr[rt][e]};
          const double big = acc[2 * e] * 65536.0;  // exact (power-of-two scale)
          const double k = __builtin_rint(big * mf.qinv);
          const double r = reduce_f(__builtin_fma(-k, mf.q, big) + acc[2 * e + 1], mf);
          edst[rt][e][pos] = (uint32_t)(int32_t)r;
        }
      }
    }
  }
}

// small instances (D <= 32, S <= 12: every ModUp and ModDown of the 24+8
// configuration) and the general bound
template <int RT, int KS>
__global__ __launch_bounds__(kThreads) void bconv_mfma_kernel(const BconvLaunch L,
                                                              const PrimeTab P) {
  bconv_body<RT, KS>(L.inst[blockIdx.y], P, blockIdx.x);
}

template <int RT, int KS>
__global__ __launch_bounds__(kThreads) void bconv_mfma_table_kernel(
    const BconvInst *__restrict__ insts, const PrimeTab P) {
  typedef const __attribute__((address_space(4))) BconvInst CInst;
  CInst &I = ((CInst *)insts)[blockIdx.y];
  bconv_body<RT, KS>(I, P, blockIdx.x);
}

}  // namespace

void launch_bconv_mfma(const Context &ctx, const BconvLaunch &L, hipStream_t st) {
  if (L.n <= 0) return;
  int dsts = 0;
  bool small = true;
  for (int i = 0; i < L.n; i++) {
    dsts += L.inst[i].D;
    small = small && L.inst[i].D <= 32 && L.inst[i].S <= 12;
  }
  if (small)
    launch(ctx, "bconv_mfma", dsts, bconv_mfma_kernel<4, 3>, dim3(kN / kChunk, L.n),
           dim3(kThreads), st, L, ctx.ptab);
  else
    launch(ctx, "bconv_mfma", dsts, bconv_mfma_kernel<kMaxMfmaRowTiles, kMaxMfmaKSteps>,
           dim3(kN / kChunk, L.n), dim3(kThreads), st, L, ctx.ptab);
}

void launch_bconv_mfma_table(const Context &ctx, const BconvInst *insts, int n,
                             int dsts, bool small, hipStream_t st) {
  if (n <= 0) return;
  if (small)
    launch(ctx, "bconv_mfma[batch]", dsts, bconv_mfma_table_kernel<4, 3>,
           dim3(kN / kChunk, n), dim3(kThreads), st, insts, ctx.ptab);
  else
    launch(ctx, "bconv_mfma[batch]", dsts,
           bconv_mfma_table_kernel<kMaxMfmaRowTiles, kMaxMfmaKSteps>,
           dim3(kN / kChunk, n), dim3(kThreads), st, insts, ctx.ptab);
}

}  // namespace ck
