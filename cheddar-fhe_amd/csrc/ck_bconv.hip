// Fast base conversion (BaseConv), the dense product behind ModUp and
// ModDown (reference kernel ModSwitchMatrixMult, src/core/ModSwitch.cu:23-171).
//
//   dst[m][k] = ReduceMontgomery( sum_i src[i][k] * T[m][i] )  mod t_m
//
// src holds s centred residues per coefficient (|v| <= (q-1)/2), T[m][i] =
// centred (prod_{j != i} s_j) * R^2 mod t_m, so dst = X * R mod t_m with
// X = sum_i v_i * s^_i, an exact integer.
//
// MI355X mapping: one thread per coefficient and a block of up to kDstBlock
// destination limbs. The table row of a destination limb is uniform across
// the workgroup (scalar loads), the s source values live in VGPRs, products
// accumulate in int64 and are folded by t_m * 2^32 every four terms so the
// final Montgomery reduction sees |acc| <= t_m * 2^31.
#include "ck_internal.h"

namespace ck {
namespace {

constexpr int kThreads = 256;
constexpr int kDstBlock = 8;

__device__ __forceinline__ int64_t fold(int64_t acc, uint32_t t) {
  const int64_t t32 = (int64_t)t << 32;
  acc += (acc < 0) ? t32 : 0;
  acc -= (acc >= (t32 >> 1)) ? t32 : 0;
  return acc;
}

template <int S>
__global__ __launch_bounds__(kThreads) void bconv_kernel(const BconvLaunch B,
                                                         const PrimeTab P) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  const int m0 = blockIdx.y * kDstBlock;
  const int s = (S > 0) ? S : B.s;
  int32_t v[kMaxBconvSrc];
#pragma unroll
  for (int i = 0; i < kMaxBconvSrc; i++)
    if (i < s) v[i] = B.src[blockIdx.z][(size_t)i * 65536 + k];
  for (int mm = 0; mm < kDstBlock; mm++) {
    const int m = m0 + mm;
    if (m >= B.d) break;
    const int p = B.dst_prime[m];
    const uint32_t t = P.q[p];
    const int32_t tinv = P.qinv[p];
    const int32_t *__restrict__ row = B.table + m * s;
    int64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kMaxBconvSrc; i++) {
      if (i < s) {
        acc += (int64_t)v[i] * (int64_t)row[i];
        if ((i & 3) == 3) acc = fold(acc, t);
      }
    }
    acc = fold(acc, t);
    B.dst[blockIdx.z][(size_t)B.dst_limb[m] * 65536 + k] = mont_reduce(acc, t, tinv);
  }
}

}  // namespace

void launch_bconv(const Context &ctx, const BconvLaunch &B, int npoly,
                  hipStream_t st) {
  if (B.d <= 0 || B.s <= 0 || npoly <= 0) return;
  dim3 grid(ctx.n / kThreads, (B.d + kDstBlock - 1) / kDstBlock, npoly);
  switch (B.s) {
    case 8:
      hipLaunchKernelGGL(bconv_kernel<8>, grid, dim3(kThreads), 0, st, B,
                         ctx.ptab);
      break;
    case 9:
      hipLaunchKernelGGL(bconv_kernel<9>, grid, dim3(kThreads), 0, st, B,
                         ctx.ptab);
      break;
    default:
      hipLaunchKernelGGL(bconv_kernel<0>, grid, dim3(kThreads), 0, st, B,
                         ctx.ptab);
      break;
  }
}

}  // namespace ck
