// BaseConv on the FP64 matrix cores (see ck_bconv.hip).
#pragma once

#include "ck_internal.h"

namespace ck {

constexpr int kMaxMfmaDst = 64;       // destination limbs of one instance
constexpr int kMaxMfmaRowTiles = kMaxMfmaDst / 8;
constexpr int kMaxMfmaKSteps = kMaxBconvSrc / 4;
constexpr int kMaxBconvInst = 4;      // instances per kernel-argument launch

// One BaseConv: S contiguous centred source limbs -> D destination limbs
// (limb dst_limb[d] of dst, prime dst_prime[d]); tab holds D x S x {hi, lo}
// (the ModUp group's or ModDown plan's FP64 table).
struct BconvInst {
  const int32_t *src;
  uint32_t *dst;
  const double *tab;
  int S, D;
  int canonical;  // 1: outputs in [0, t); 0: lazy int32 for a transform pass
  uint16_t dst_limb[kMaxMfmaDst];
  uint8_t dst_prime[kMaxMfmaDst];
};

struct BconvLaunch {
  int n;
  BconvInst inst[kMaxBconvInst];
};

void launch_bconv_mfma(const Context &ctx, const BconvLaunch &L, hipStream_t st);
// table form: n instances in device memory; dsts = total destination limbs,
// max_s / max_d = the largest S and D among the instances
void launch_bconv_mfma_table(const Context &ctx, const BconvInst *insts, int n,
                             int dsts, int max_s, int max_d, hipStream_t st);

}  // namespace ck
