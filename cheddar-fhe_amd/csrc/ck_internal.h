// Internal declarations of the MI355X CKKS key-switching core.
//
// Not part of the C ABI (include/cheddar_amd.h); shared by the HIP kernels
// (ck_*.hip) and the host orchestration (ck_context.cpp, ck_ops.cpp).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "ck_arith.h"

namespace ck {

// ---- status codes (mirrored in include/cheddar_amd.h) ----
enum Status : int {
  kOk = 0,
  kErrInvalidArg = 1,
  kErrUnsupported = 2,
  kErrHip = 3,
  kErrOutOfMemory = 4,
};

// Device view of the per-prime constants (global prime order
// [terminal reversed | main | auxiliary]).
struct PrimeTab {
  const uint32_t *q;
  const int32_t *qinv;  // q^-1 mod 2^32 (signed)
  const double *qd;     // q as double
  const double *qr;     // 1/q
};

// Per-launch limb plan, passed by value as a kernel argument so that no
// host-to-device copy happens inside an operation (graph-capture safe).
constexpr int kMaxLaunchLimbs = 112;

enum NttFlags : int {
  kNttPreMul = 1,           // forward: x *= c1 before the transform
  kNttModDownEpilogue = 2,  // forward: out = (src2 * c2 - NTT(x)) * c3
  kNttEpiloguePad = 4,      // forward: apply c2 in the epilogue
  kInttCentred = 8,         // inverse: centred int32 output
  kInttLoadProduct = 16,    // inverse: input = src (*) src2 (Montgomery)
  kNttBconvInput = 32,      // forward: input = BaseConv of bsrc (see below)
  kNttColsOnly = 64,        // forward: column stages only (rows fused later)
};

// Per-limb fields of a kernel-argument plan, kept together in one record
// (24 bytes, dword aligned): hipcc for gfx950 has folded a byte-indexed base
// into the scalar load of a separate dword array (s_load with a base that
// is not dword aligned, whose low bits the hardware drops), reading the
// constant of the neighbouring limb.
struct NttLimb {
  uint32_t c1, c2, c3;  // plain-form constants (already carry R^-1)
  uint16_t src_limb, dst_limb;
  int16_t src2_limb;
  uint16_t btab_off;
  uint8_t prime, sel;
  uint16_t pad_;
};
static_assert(sizeof(NttLimb) == 24, "NttLimb is one 24-byte record");

struct NttLaunch {
  // up to four base pointers per role; limb i uses src[sel & 3],
  // dst[(sel >> 2) & 3], src2[(sel >> 4) & 3], bsrc[(sel >> 6) & 3]
  const uint32_t *src[4];
  uint32_t *dst[4];
  const uint32_t *src2[4];
  // kNttBconvInput: the transform input of limb i is
  //   sum_j bsrc[j] * T_j  mod q, T_j = B[btab_off + 2j] * 2^16 +
  //   B[btab_off + 2j + 1], B = btab[(sel >> 6) & 3]
  // over bs[(sel >> 6) & 3] contiguous centred int32 source limbs
  // (BaseConv, exact)
  const int32_t *bsrc[4];
  const double *btab[4];
  int bs[4];
  int n;
  int flags;
  NttLimb limb[kMaxLaunchLimbs];
};

// One limb of a transform pass with every pointer resolved; the form a
// batched plan keeps in device memory (one table per pass, 64 bytes a limb).
// src2 == nullptr: no product operand / no kept limb; bsrc == nullptr: no
// BaseConv input. flags as NttLaunch::flags.
struct LimbJob {
  const uint32_t *src;
  uint32_t *dst;
  const uint32_t *src2;
  const int32_t *bsrc;
  const double *btab;  // this output limb's row: bs x {hi, lo}
  uint32_t c1, c2, c3;
  uint16_t prime, bs;
  int32_t flags;
  int32_t pad_;
};
static_assert(sizeof(LimbJob) == 64, "LimbJob is one 64-byte record");

// source limbs of one BaseConv: up to 5 terminal + 12 auxiliary primes in
// the reference's parameter files (bootparam_35.json); 24 keeps every exact
// FP64 and int8 sum within its stated bound
constexpr int kMaxBconvSrc = 24;
constexpr int kMaxBconvDst = kMaxLaunchLimbs;

// Tables of one ModDown flavour (ModDown, Rescale, ModDownAndRescale) at one
// level, mirroring ModSwitchHandler's constants (ModSwitch.cu:233-411).
struct ModDownPlan {
  bool valid = false;
  int src_len = 0, dst_len = 0;
  std::vector<int> src_limb;   // limb index in the source poly
  std::vector<int> src_prime;  // global prime index
  std::vector<uint32_t> c1;    // INTT post-constant (plain form)
  std::vector<int> dst_prime;  // global prime index of each output limb
  std::vector<int> src2_limb;  // source limb added in the epilogue, -1 none
  std::vector<uint32_t> pad;   // epilogue padding (plain)
  std::vector<uint32_t> inv_prod;  // epilogue (prod src)^-1 (plain)
  bool use_pad = false;
  double *d_ftab = nullptr;    // dst_len x src_len x {hi, lo}, see NttLaunch
};

struct ModUpGroup {
  int s0 = 0, s1 = 0;           // own limb range [s0, s1) within the level
  std::vector<int> dst_limb;    // limb index in the (num_q + alpha) output
  std::vector<int> dst_prime;   // global prime
  double *d_ftab = nullptr;     // dst x src x {hi, lo}
};

struct LevelPlan {
  int nm = 0, nt = 0, num_q = 0, beta = 0, pad = 0;
  std::vector<int> q_prime;        // global prime of each q limb
  std::vector<uint32_t> modup_c1;  // per q limb (plain form)
  std::vector<ModUpGroup> groups;  // beta entries (empty ones have s1 <= s0)
  ModDownPlan moddown, rescale, moddown_rescale;
};

struct Context {
  int log_n = 0, n = 0;
  int max_main = 0, max_ter = 0, alpha = 0, L = 0, dnum = 0, max_level = 0;
  int num_primes = 0;
  std::vector<int> lv_main, lv_ter;
  std::vector<uint32_t> primes;
  std::vector<int32_t> qinv;
  std::vector<uint32_t> p_prod_mont;  // P mod q_i * R (q primes)
  std::vector<uint32_t> galois;       // 5^i mod 2N, i in [0, N/2]
  std::vector<uint32_t> psi;          // the 2N-th root of each prime's NTT

  // device tables
  uint32_t *d_q = nullptr;
  int32_t *d_qinv = nullptr;
  double *d_qd = nullptr, *d_qr = nullptr;
  uint32_t *d_tw_fwd = nullptr;  // psi^brv(j), plain, per prime
  uint32_t *d_tw_inv = nullptr;  // psi^-brv(j), plain, per prime
  uint32_t *d_p_prod = nullptr;  // P mod q * R, indexed by global q prime
  uint32_t *d_imag = nullptr;    // psi^(N/2) * R (X^(N/2) in NTT form), per prime
  PrimeTab ptab{};
  // int8 matrix-core transform tables (built on first use, ck_i8.hip)
  const int32_t *i8_planes = nullptr, *i8_corr = nullptr;
  const uint32_t *i8_tw = nullptr;

  std::vector<LevelPlan> levels;
  // Key switching at level -1 (the ModSwitchHandler of level -1 with
  // Parameter::GetSSENumAux auxiliary primes, ModSwitch.cu:175-232): the
  // short base's q limbs form one ModUp group, the first short_naux
  // auxiliary primes are its P', and d_p_prod_short holds P' mod q * R by
  // global prime (Context.cpp:109-123). short_naux = 0: not set up.
  LevelPlan short_base;
  int short_naux = 0;
  uint32_t *d_p_prod_short = nullptr;

  // BaseConv as its own step on the matrix cores instead of fused into the
  // forward column pass: 1 FP64 matrix cores everywhere, 2 int8 matrix
  // cores everywhere, 0 fused everywhere, -1 (default) int8 matrix cores in
  // batched plans and fused in single operations (measured fastest)
  int bconv_mfma = -1;
  // int8 operand planes of the BaseConv tables, by FP64 table (built on
  // first use, freed with the context)
  mutable std::vector<std::pair<const double *, int64_t *>> i8_bconv_planes;
  // batched plans: build both BaseConv forms, time them, keep the faster
  bool plan_autotune = false;
  // key switching: the ModUp's forward row stages fused with the key inner
  // product (one pass, no NTT-form ModUp output) instead of a row pass and
  // a separate key_accum
  bool fuse_key_rows = true;
  // forms of ck_plan_ntt (forward) and ck_plan_bconv: see CK_OPT_PLAN_NTT_FORM
  // and CK_OPT_PLAN_BCONV_FORM in cheddar_amd.h (-1: autotune)
  int plan_ntt_form = 1;    // int8 matrix cores (measured faster)
  int plan_bconv_form = 1;  // int8 matrix cores (measured faster)
  int ntt8_blocks = 16;  // int8 transform: 16-column blocks per workgroup
  int ntt8_chunk = 0;    // int8 transform: limbs per column/row chunk (0: all)

  // workspace for multi-kernel operations (one operation at a time per
  // context; concurrent calls on different streams race)
  uint32_t *ws = nullptr;
  size_t ws_words = 0;

  std::vector<void *> allocations;

  // optional per-kernel timing: HIP events recorded around every launch on
  // the caller's stream while enabled (ck_profile_*)
  struct ProfEvent {
    int kernel;
    hipEvent_t start, stop;
  };
  mutable bool prof_on = false;
  mutable std::vector<ProfEvent> prof_events;
  mutable std::vector<hipEvent_t> prof_pool;
  mutable std::vector<std::string> prof_names;
};

// Per-kernel timing: while profiling is on, prof_events hands out an event
// pair for the next launch of `kernel` over `limbs` limbs, and launch()
// passes it to hipExtLaunchKernelGGL, which stamps the dispatch's own start
// and end into the events (no extra packets between kernels).
bool prof_events(const Context &c, const char *kernel, int limbs,
                 hipEvent_t *start, hipEvent_t *stop);

template <typename F, typename... Args>
inline void launch(const Context &c, const char *name, int limbs, F kernel,
                   dim3 grid, dim3 block, hipStream_t st, Args... args) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (prof_events(c, name, limbs, &e0, &e1))
    hipExtLaunchKernelGGL(kernel, grid, block, 0, st, e0, e1, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
}
int prof_report(const Context &c, std::vector<std::string> &names,
                std::vector<double> &total_ms, std::vector<long> &calls);
void prof_reset(const Context &c);

// ---- context.cpp ----
int context_create(int log_n, const uint32_t *main_primes, int n_main,
                   const uint32_t *ter_primes, int n_ter,
                   const uint32_t *aux_primes, int n_aux,
                   const int *level_cfg, int n_levels, Context **out);
void context_destroy(Context *ctx);
void set_error(const std::string &msg);
const char *last_error();
int check_hip(hipError_t e, const char *what);
int global_prime(const Context &c, int num_main, int num_ter, int limb);
// level -1: the short base of num_main + num_ter primes (the bottom level's
// primes less the additional base) with as many auxiliary primes
int build_short_base(Context &c, int num_main, int num_ter);
// the key-switching plan and auxiliary prime count of a level (-1 = the
// short base when built)
inline bool has_switch_level(const Context &c, int level) {
  return (level >= 0 && level <= c.max_level) || (level == -1 && c.short_naux > 0);
}
inline const LevelPlan &level_plan(const Context &c, int level) {
  return level < 0 ? c.short_base : c.levels[level];
}
inline int level_naux(const Context &c, int level) {
  return level < 0 ? c.short_naux : c.alpha;
}

// host number theory (independent restatement of PrimeUtils.h)
uint32_t h_mulmod(uint32_t a, uint32_t b, uint32_t q);
uint32_t h_powmod(uint32_t a, uint64_t e, uint32_t q);
uint32_t h_invmod(uint32_t a, uint32_t q);
uint32_t h_to_mont(uint32_t a, uint32_t q);

// BaseConv constants from `src` to `dst` (global prime indices), see
// ck_context.cpp: c1 = the INTT post-constant per source, table = the
// reference's centred (prod_{j!=i} s_j) R^2, ftab = (prod_{j!=i} s_j) R split
// into {hi, lo} 16-bit halves for the exact FP64 forms
void bconv_constants(const Context &c, const std::vector<int> &src,
                     const std::vector<int> &dst,
                     const std::vector<int> &restore,
                     std::vector<uint32_t> &c1, std::vector<int32_t> &table,
                     std::vector<double> &ftab);

// ---- launchers ----
void launch_ntt_fwd(const Context &ctx, const NttLaunch &L, hipStream_t st);
void launch_intt(const Context &ctx, const NttLaunch &L, hipStream_t st);
// table form: n jobs in device memory; flags_any = OR of the jobs' flags
// (names the launch for profiling); max_bs = the largest BaseConv source
// count among the jobs (selects the column kernel's instantiation)
void launch_ntt_fwd_table(const Context &ctx, const LimbJob *jobs, int n,
                          int flags_any, int max_bs, hipStream_t st);
void launch_intt_table(const Context &ctx, const LimbJob *jobs, int n,
                       int flags_any, hipStream_t st);

}  // namespace ck
