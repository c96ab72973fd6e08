// Element-wise launch descriptors (see ck_elem.hip).
#pragma once

#include "ck_internal.h"

namespace ck {

constexpr int kMaxPolys = 3;
constexpr int kMaxDnum = 16;

// Limb -> prime map of an NP (num_main, num_ter, num_aux): limb y < num_q
// uses global prime q_prime0 + y, aux limb y uses aux_prime0 + (y - num_q).
struct EwNp {
  int num_q;
  int num_total;
  int q_prime0;
  int aux_prime0;
};

// Up to kMaxPolys polynomials sharing one layout; aux limbs sit `gap` limbs
// after the q part (the reference's DvView extra_, PtrList.h:37-63).
struct EwOperand {
  uint32_t *ptr[kMaxPolys];
  int gap;
};

struct KeyAccumArgs {
  const uint32_t *modup;  // beta blocks of num_total limbs
  const uint32_t *key_b[kMaxDnum];  // key blocks (limb layout: key_q0, key_aux0)
  const uint32_t *key_a[kMaxDnum];
  // key block limb of the first q limb and of the first aux limb (full-size
  // blocks: the global primes max_ter - nt and L; the short-base key: 0 and
  // num_q)
  int key_q0, key_aux0;
  int group[kMaxDnum];
  int ngroups;
  // ModUp source (NTT domain, num_q limbs): a group's own limbs [own0, own1)
  // are read from here instead of the ModUp output (ModUp copies them
  // unchanged, ModSwitch.cu:466-470)
  const uint32_t *x;
  int own0[kMaxDnum];
  int own1[kMaxDnum];
  const uint32_t *add_b;  // num_q limbs or null
  const uint32_t *add_a;
  // HMult form: tensor = (b1, a1, b2, a2); the key-switched polynomial is
  // d2 = a1 (*) a2 (own limbs formed here) and the fold adds P * d0,
  // P * d1 with d0 = b1 (*) b2, d1 = b1 (*) a2 + a1 (*) b2
  const uint32_t *tensor[4];
  int use_tensor;
  const uint32_t *p_prod;  // by global prime
  uint32_t *out_b;
  uint32_t *out_a;
};

// Hoisted baby step (HoistHandler::BSFusedKeyMult + BSFusedKernel,
// src/extension/Hoist.cu:21-78, :207-327) for up to kMaxBsRot rotations of
// one ModUp'd polynomial:
//   acc_r[y][x] = sum_g modup_g[y][x] (*) key_{r,g}[prime(y)][x]
//                 (+ P (*) b[y][x] on q limbs)
// stored at the automorphism's position (scatter by g_r^-1 = 5^(N/2 - r)).
constexpr int kMaxBsRot = 8;
constexpr int kMaxBsGroups = 8;

struct BsFusedArgs {
  const uint32_t *modup;  // beta blocks of num_total limbs
  const uint32_t *x;      // ModUp source (own limbs are read from here)
  const uint32_t *b;      // pseudo-ModUp input (num_q limbs)
  const uint32_t *p_prod; // by global prime
  int ngroups;
  int nrot;
  int group[kMaxBsGroups];
  int own0[kMaxBsGroups];
  int own1[kMaxBsGroups];
  const uint32_t *key_b[kMaxBsRot][kMaxBsGroups];
  const uint32_t *key_a[kMaxBsRot][kMaxBsGroups];
  uint32_t *out_b[kMaxBsRot];  // num_total limbs each
  uint32_t *out_a[kMaxBsRot];
  uint32_t ginv[kMaxBsRot];
};

// Sum of products over up to kMaxAccTerms terms for npoly (1..2)
// polynomials sharing the multiplier of each term (PAccum / CAccum,
// ElementWise.cu:225-366, and the giant-step GSFusedKernel,
// src/extension/Hoist.cu:80-118):
//   dst[p] = (accumulate ? dst[p] : 0) + sum_t x[t][p] (*) y[t]
// where y[t] is a polynomial (per coefficient) or, with y_const, one value
// per limb. All operands hold num_total contiguous limbs.
constexpr int kMaxAccTerms = 16;

// Giant-step accumulation of a hoisted linear transform (GSFusedKernel,
// src/extension/Hoist.cu:80-118): for each of ngs giant steps g,
//   dst[g][j] = sum_t x[t][j] (*) pt[g][t]   (j = b, a; pt null: no term)
// with every baby step x[t] read once for all giant steps.
constexpr int kMaxGsBaby = 16;

// Sum of automorphisms (PermuteAccum, src/core/ElementWise.cu:394-501):
//   dst[j] = (accumulate ? dst[j] : 0) + sum_s sigma_{g_s}(src[s][j])
// for j < npoly (1..2); a null src[s][j] adds nothing; g = 1 is the identity.
constexpr int kMaxPermAcc = 8;
struct PermAccArgs {
  const uint32_t *src[kMaxPermAcc][2];
  uint32_t g[kMaxPermAcc];
  uint32_t *dst[2];
  int nsrc, npoly, accumulate;
};
constexpr int kMaxGsGiant = 8;  // per launch
struct GsAccumArgs {
  const uint32_t *x[kMaxGsBaby][2];
  const uint32_t *pt[kMaxGsGiant][kMaxGsBaby];
  uint32_t *dst[kMaxGsGiant][2];
  int nbs, ngs;
};

struct AccumArgs {
  const uint32_t *x[kMaxAccTerms][2];
  const uint32_t *y[kMaxAccTerms];
  uint32_t *dst[2];
  int nterms;
  int npoly;
  int y_const;
  int accumulate;
};

// Batched (table) forms: item i of a plan stage has its own layout and
// arguments; job j covers limb jobs[j].y of item jobs[j].item.
struct EwJob {
  uint16_t item;
  uint16_t y;
};
struct KeyAccumItem {
  EwNp np;
  KeyAccumArgs A;
};
struct BsFusedItem {
  EwNp np;
  BsFusedArgs A;
};

void launch_binary(const Context &ctx, int op, const EwNp &np,
                   const EwOperand &x, const EwOperand &y, const EwOperand &out,
                   int npoly, hipStream_t st);
void launch_unary(const Context &ctx, int op, const EwNp &np,
                  const EwOperand &x, const EwOperand &out, int npoly,
                  hipStream_t st);
void launch_const(const Context &ctx, int op, const EwNp &np,
                  const EwOperand &x, const uint32_t *consts, int const_gap,
                  const EwOperand &out, int npoly, hipStream_t st);
void launch_imag_unit(const Context &ctx, const EwNp &np, const EwOperand &x,
                      const EwOperand &out, int npoly, hipStream_t st);
// Exact CRT lift of one- or two-prime coefficient-domain residues to up to
// kMaxLiftDst destination primes (global prime indices).
constexpr int kMaxLiftDst = 128;
struct LiftArgs {
  int nsrc;
  uint32_t q0, q1;
  uint16_t dst_prime[kMaxLiftDst];
};
void launch_crt_lift(const Context &ctx, const uint32_t *src, const LiftArgs &A,
                     int ndst, uint32_t *dst, hipStream_t st);
void launch_accum(const Context &ctx, const EwNp &np, const AccumArgs &A,
                  hipStream_t st);
void launch_gs_accum(const Context &ctx, const EwNp &np, const GsAccumArgs &A,
                     hipStream_t st);
void launch_permute_accum(const Context &ctx, const EwNp &np, const PermAccArgs &A,
                          hipStream_t st);
// bytes from src to dst (ranges equal or disjoint) on the device: 16-byte
// vectors when both are 16-byte aligned and bytes % 16 == 0, else
// hipMemcpyAsync; returns a status code
int launch_copy(const Context &ctx, void *dst, const void *src, size_t bytes,
                hipStream_t st);
void launch_tensor(const Context &ctx, const EwNp &np, const EwOperand &c1,
                   const EwOperand &c2, const EwOperand &out, hipStream_t st);
void launch_key_accum(const Context &ctx, const EwNp &np,
                      const KeyAccumArgs &A, hipStream_t st);
void launch_permute(const Context &ctx, const EwNp &np, const EwOperand &x,
                    const EwOperand &out, uint32_t g, int npoly,
                    hipStream_t st);
void launch_bs_fused(const Context &ctx, const EwNp &np, const BsFusedArgs &A,
                     hipStream_t st);
void launch_key_accum_table(const Context &ctx, const KeyAccumItem *items,
                            const EwJob *jobs, int njobs, hipStream_t st);
void launch_bs_fused_table(const Context &ctx, const BsFusedItem *items,
                           const EwJob *jobs, int njobs, hipStream_t st);
// forward row stages of a ModUp fused with the key inner product (the
// column pass ran with kNttColsOnly; A.modup holds its lazy int32 output);
// writes A.out_b / A.out_a exactly as launch_key_accum would (ck_ntt.hip)
void launch_key_rows(const Context &ctx, const EwNp &np, const KeyAccumArgs &A,
                     hipStream_t st);
void launch_key_rows_table(const Context &ctx, const KeyAccumItem *items,
                           const EwJob *jobs, int njobs, hipStream_t st);

inline EwNp make_np(const Context &c, int num_main, int num_ter, int num_aux) {
  EwNp np;
  np.num_q = num_main + num_ter;
  np.num_total = np.num_q + num_aux;
  np.q_prime0 = c.max_ter - num_ter;
  np.aux_prime0 = c.L;
  return np;
}

}  // namespace ck
