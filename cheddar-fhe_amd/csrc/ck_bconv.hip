// BaseConv on the FP64 matrix cores (ModSwitchMatrixMult,
// src/core/ModSwitch.cu:23-171): for every coefficient position x and
// destination limb d,
//   y_d(x) = sum_k v_k(x) * T_dk  mod t_d,   T_dk = hi * 2^16 + lo,
// over S centred int32 source limbs v_k (|v| < 2^30). As a matrix product
// the rows are (destination, half) pairs, the reduction runs over the
// sources and the columns over positions; with 16-bit table halves every
// product is below 2^46 and every sum of S <= 24 products below 2^51, so
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
// the forward column pass reads, or canonical residues (standalone BaseConv).
#include "ck_bconv.h"

#include <algorithm>

namespace ck {
namespace {

constexpr int kN = 65536;
constexpr int kThreads = 256;
constexpr int kChunk = 512;  // positions per workgroup (128 per wave)

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kTiles = kChunk / 4 / 16;  // 16-position tiles per wave

// I: the instance, read in place (kernel argument or constant-space table);
// RT: static bound on the row tiles (8 destinations each), KS: the k-steps
// (4 sources each) every instance of the launch is padded to, so that the
// operand arrays stay in registers. Every load is unconditional (an index
// past the instance's S or D is clamped to 0 and its value replaced by 0
// afterwards): a load inside a branch is waited for before the branch ends,
// which would serialise the source loads of a wave.
template <int RT, int KS, class Inst>
__device__ __forceinline__ void bconv_body(Inst &I, const PrimeTab &P, int chunk) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4, n16 = lane & 15;
  const int S = I.S, D = I.D;
  const int nrt = (D + 7) >> 3;
  const int base = chunk * kChunk + wave * (kChunk / 4);

  // every source value this lane needs, loaded up front (one memory
  // latency per wave): source k = 4 ks + l / 16 at position base + 16 t + l % 16
  int32_t raw[kTiles][KS];
#pragma unroll
  for (int ks = 0; ks < KS; ks++) {
    const int k = ks * 4 + g;
    const int32_t *sp = I.src + (size_t)(k < S ? k : 0) * kN + base + n16;
#pragma unroll
    for (int t = 0; t < kTiles; t++) raw[t][ks] = sp[t * 16];
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
      const bool ok = d < D && k < S;
      const double v = I.tab[(ok ? d * S + k : 0) * 2 + half];
      a[rt][ks] = ok ? v : 0.0;
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
      const bool ok = d < D;
      const int dd = ok ? d : 0;
      const int p = I.dst_prime[dd];
      const int limb = I.dst_limb[dd];
      eq[rt][e] = P.qd[p];
      eqr[rt][e] = P.qr[p];
      edst[rt][e] = ok ? I.dst + (size_t)limb * kN : nullptr;
    }

  // sources past S have zero table entries (a = 0), so their k-step
  // products vanish and every instance runs the same KS steps
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
      for (int ks = 0; ks < KS; ks++)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt][ks], b[ks], acc, 0, 0, 0);
#pragma unroll
      for (int e = 0; e < 2; e++) {
        if (edst[rt][e]) {
          const ModF mf{eq[rt][e], eqr[rt][e]};
          const double big = acc[2 * e] * 65536.0;  // exact (power-of-two scale)
          const double k = __builtin_rint(big * mf.qinv);
          double r = reduce_f(__builtin_fma(-k, mf.q, big) + acc[2 * e + 1], mf);
          if (I.canonical) r = r < 0.0 ? r + mf.q : r;
          edst[rt][e][pos] = (uint32_t)(int32_t)r;
        }
      }
    }
  }
}

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

// the instantiation for a launch whose instances have at most max_s sources
// and max_d destinations: row tiles 4 (D <= 32) or 8, k-steps 2, 3 or 4
template <int RT, class F>
void pick_ks(int max_s, F &&f) {
  const int ks = (max_s + 3) >> 2;
  if (ks <= 2)
    f(bconv_mfma_kernel<RT, 2>, bconv_mfma_table_kernel<RT, 2>);
  else if (ks == 3)
    f(bconv_mfma_kernel<RT, 3>, bconv_mfma_table_kernel<RT, 3>);
  else if (ks == 4)
    f(bconv_mfma_kernel<RT, 4>, bconv_mfma_table_kernel<RT, 4>);
  else
    f(bconv_mfma_kernel<RT, kMaxMfmaKSteps>, bconv_mfma_table_kernel<RT, kMaxMfmaKSteps>);
}

template <class F>
void pick(int max_s, int max_d, F &&f) {
  if (max_d <= 32)
    pick_ks<4>(max_s, f);
  else
    pick_ks<kMaxMfmaRowTiles>(max_s, f);
}

}  // namespace

void launch_bconv_mfma(const Context &ctx, const BconvLaunch &L, hipStream_t st) {
  if (L.n <= 0) return;
  int dsts = 0, max_s = 1, max_d = 1;
  for (int i = 0; i < L.n; i++) {
    dsts += L.inst[i].D;
    max_s = std::max(max_s, L.inst[i].S);
    max_d = std::max(max_d, L.inst[i].D);
  }
  pick(max_s, max_d, [&](auto kern, auto) {
    launch(ctx, "bconv_mfma", dsts, kern, dim3(kN / kChunk, L.n), dim3(kThreads), st, L,
           ctx.ptab);
  });
}

void launch_bconv_mfma_table(const Context &ctx, const BconvInst *insts, int n,
                             int dsts, int max_s, int max_d, hipStream_t st) {
  if (n <= 0) return;
  pick(max_s, max_d, [&](auto, auto kern) {
    launch(ctx, "bconv_mfma[batch]", dsts, kern, dim3(kN / kChunk, n), dim3(kThreads), st,
           insts, ctx.ptab);
  });
}

}  // namespace ck
