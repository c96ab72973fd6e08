// Host orchestration of the key-switching path on one HIP stream:
// ModUp, ModDown / Rescale / ModDownAndRescale, the key inner product and
// the evaluator compositions HMult, Relinearize, MultKey, HRot, HConj,
// Rescale (reference: src/core/ModSwitch.cu:413-684,
// src/core/Context.cpp:316-871).
//
// Every launch plan is built on the host and passed by value; temporaries
// come from the context workspace, so an operation issues no allocation, no
// host<->device copy and no synchronisation.
//
// Fusions relative to the reference's kernel sequence (results identical):
//   * BaseConv runs inside the forward NTT's column pass (ModUp and ModDown),
//   * HMult's Tensor product a1 (*) a2 is formed while the ModUp INTT loads,
//     and d0 = b1 (*) b2, d1 = b1 (*) a2 + a1 (*) b2 inside the key inner
//     product's P * (d0, d1) fold,
//   * ModUp's own-group limbs are read from the source by the inner product
//     instead of being copied.
#include "ck_ops.h"

#include <algorithm>
#include <cstring>

namespace ck {

namespace {

constexpr size_t kLimb = 65536;

int launch_status(const char *what) {
  return check_hip(hipGetLastError(), what);
}

struct NttJob {
  int prime, sel, src_limb, dst_limb, src2_limb;
  uint32_t c1, c2, c3;
  int btab_off;
};

// Split a list of limb jobs into launches of at most kMaxLaunchLimbs.
void run_jobs(const Context &c, bool forward, const NttLaunch &base,
              const std::vector<NttJob> &jobs, hipStream_t st) {
  for (size_t off = 0; off < jobs.size(); off += kMaxLaunchLimbs) {
    NttLaunch L = base;
    L.n = (int)std::min<size_t>(kMaxLaunchLimbs, jobs.size() - off);
    for (int i = 0; i < L.n; i++) {
      const NttJob &j = jobs[off + i];
      L.prime[i] = (uint8_t)j.prime;
      L.sel[i] = (uint8_t)j.sel;
      L.src_limb[i] = (uint16_t)j.src_limb;
      L.dst_limb[i] = (uint16_t)j.dst_limb;
      L.src2_limb[i] = (int16_t)j.src2_limb;
      L.btab_off[i] = (uint16_t)j.btab_off;
      L.c1[i] = j.c1;
      L.c2[i] = j.c2;
      L.c3[i] = j.c3;
    }
    if (forward)
      launch_ntt_fwd(c, L, st);
    else
      launch_intt(c, L, st);
  }
}

NttLaunch empty_launch() {
  NttLaunch L;
  std::memset(&L, 0, sizeof(L));
  return L;
}

uint32_t r_inv_mod(uint32_t q) { return h_invmod(h_to_mont(1, q), q); }

}  // namespace

// ---------------------------------------------------------------------------
// NTTHandler
// ---------------------------------------------------------------------------

int op_ntt(const Context &c, uint32_t *dst, const uint32_t *src, int nm,
           int nt, int na, int src_gap, bool mont_conv, hipStream_t st) {
  NttLaunch L = empty_launch();
  L.src[0] = src;
  L.dst[0] = dst;
  L.flags = mont_conv ? kNttPreMul : 0;
  const int num_q = nm + nt;
  std::vector<NttJob> jobs;
  for (int t = 0; t < num_q + na; t++) {
    const int p = global_prime(c, nm, nt, t);
    // NTT(..., montgomery_conversion): x * R^2 * R^-1 = x * R
    const uint32_t cv = mont_conv ? h_to_mont(1, c.primes[p]) : 1u;
    jobs.push_back({p, 0, t < num_q ? t : t + src_gap, t, -1, cv, 0, 0, 0});
  }
  run_jobs(c, true, L, jobs, st);
  return launch_status("ntt");
}

// INTT followed by x * const * R^-1 with per-limb constants given in the
// reference's convention (the value it passes to MultConst).
int op_intt_const(const Context &c, uint32_t *dst, const uint32_t *src,
                  int nm, int nt, int na, int src_gap,
                  const uint32_t *consts_mont, bool centred, hipStream_t st) {
  NttLaunch L = empty_launch();
  L.src[0] = src;
  L.dst[0] = dst;
  L.flags = centred ? kInttCentred : 0;
  const int num_q = nm + nt;
  std::vector<NttJob> jobs;
  for (int t = 0; t < num_q + na; t++) {
    const int p = global_prime(c, nm, nt, t);
    const uint32_t q = c.primes[p];
    const uint32_t cv = h_mulmod(consts_mont[t] % q, r_inv_mod(q), q);
    jobs.push_back({p, 0, t < num_q ? t : t + src_gap, t, -1, cv, 0, 0, 0});
  }
  run_jobs(c, false, L, jobs, st);
  return launch_status("intt");
}

int op_intt(const Context &c, uint32_t *dst, const uint32_t *src, int nm,
            int nt, int na, int src_gap, bool mont_conv, hipStream_t st) {
  // NTT.cu:673-741: N^-1 (leave Montgomery form) or N^-1 * R (keep it)
  std::vector<uint32_t> cs;
  for (int t = 0; t < nm + nt + na; t++) {
    const uint32_t q = c.primes[global_prime(c, nm, nt, t)];
    const uint32_t ninv = h_invmod((uint32_t)c.n % q, q);
    cs.push_back(mont_conv ? ninv : h_to_mont(ninv, q));
  }
  return op_intt_const(c, dst, src, nm, nt, na, src_gap, cs.data(), false, st);
}

// ---------------------------------------------------------------------------
// ModSwitchHandler
// ---------------------------------------------------------------------------

// ModUp (ModSwitch.cu:429-509). dst holds beta blocks of (num_q + alpha)
// limbs; v is a workspace of num_q limbs. When prod2 is given the source is
// src (*) prod2 (HMult's a1 (*) a2). copy_own = false leaves each group's own
// limbs unwritten (the fused inner product reads them from the source).
int op_modup(const Context &c, int level, uint32_t *dst, const uint32_t *src,
             uint32_t *v, hipStream_t st, bool copy_own,
             const uint32_t *prod2) {
  const LevelPlan &lp = c.levels[level];
  const int tot = lp.num_q + c.alpha;
  // 1. INTT of all q limbs with the per-group constant, centred output
  {
    NttLaunch L = empty_launch();
    L.src[0] = src;
    L.src2[0] = prod2;
    L.dst[0] = v;
    L.flags = kInttCentred | (prod2 ? kInttLoadProduct : 0);
    std::vector<NttJob> jobs;
    for (int t = 0; t < lp.num_q; t++)
      jobs.push_back({lp.q_prime[t], 0, t, t, t, lp.modup_c1[t], 0, 0, 0});
    run_jobs(c, false, L, jobs, st);
  }
  // 2. BaseConv of each group fused into the forward NTT of its outputs;
  // up to four groups share a launch (one selector each)
  NttLaunch L = empty_launch();
  L.flags = kNttBconvInput;
  std::vector<NttJob> jobs;
  int nsel = 0;
  auto flush = [&]() {
    run_jobs(c, true, L, jobs, st);
    jobs.clear();
    nsel = 0;
  };
  for (int g = 0; g < lp.beta; g++) {
    const ModUpGroup &gr = lp.groups[g];
    if (gr.s1 <= gr.s0) continue;
    const int s = gr.s1 - gr.s0;
    const int d = (int)gr.dst_limb.size();
    if (nsel == 4 || jobs.size() + d > (size_t)kMaxLaunchLimbs) flush();
    const int sel = nsel++;
    L.src[sel] = dst;
    L.dst[sel] = dst;
    L.bsrc[sel] = reinterpret_cast<const int32_t *>(v + (size_t)gr.s0 * kLimb);
    L.btab[sel] = gr.d_ftab;
    L.bs[sel] = s;
    for (int m = 0; m < d; m++) {
      const int limb = g * tot + gr.dst_limb[m];
      jobs.push_back({gr.dst_prime[m], sel | (sel << 2) | (sel << 6), limb, limb,
                      -1, 1, 0, 0, 2 * m * s});
    }
    if (copy_own) {
      int rc = check_hip(
          hipMemcpyAsync(dst + ((size_t)g * tot + gr.s0) * kLimb,
                         src + (size_t)gr.s0 * kLimb,
                         (size_t)s * kLimb * sizeof(uint32_t),
                         hipMemcpyDeviceToDevice, st),
          "modup copy");
      if (rc) return rc;
    }
  }
  if (!jobs.empty()) flush();
  return launch_status("modup");
}

const ModDownPlan *moddown_plan(const Context &c, int level, int kind) {
  if (level < 0 || level > c.max_level) return nullptr;
  const LevelPlan &lp = c.levels[level];
  const ModDownPlan *mp = kind == 0   ? &lp.moddown
                          : kind == 1 ? &lp.rescale
                                      : &lp.moddown_rescale;
  return mp->valid ? mp : nullptr;
}

// ModDownWorker (ModSwitch.cu:566-684) for npoly (<= 4) polynomials at
// once; v is a workspace of npoly * src_len limbs.
int op_moddown(const Context &c, int level, int kind, int npoly,
               uint32_t *const *dst, const uint32_t *const *src, uint32_t *v,
               hipStream_t st) {
  const ModDownPlan *mp = moddown_plan(c, level, kind);
  if (!mp) {
    set_error("no ModDown plan for this level and kind");
    return kErrInvalidArg;
  }
  if ((size_t)npoly * mp->src_len * kLimb > c.ws_words) {
    set_error("ModDown workspace too small");
    return kErrUnsupported;
  }
  // 1. INTT of the removed limbs, x const1, centred
  {
    NttLaunch L = empty_launch();
    std::vector<NttJob> jobs;
    for (int j = 0; j < npoly; j++) {
      L.src[j] = src[j];
      L.dst[j] = v + (size_t)j * mp->src_len * kLimb;
      for (int i = 0; i < mp->src_len; i++)
        jobs.push_back({mp->src_prime[i], j | (j << 2), mp->src_limb[i], i, -1,
                        mp->c1[i], 0, 0, 0});
    }
    L.flags = kInttCentred;
    run_jobs(c, false, L, jobs, st);
  }
  // 2. BaseConv fused into the forward NTT, 3. epilogue in the row pass
  {
    NttLaunch L = empty_launch();
    L.flags = kNttBconvInput | kNttModDownEpilogue |
              (mp->use_pad ? kNttEpiloguePad : 0);
    std::vector<NttJob> jobs;
    for (int j = 0; j < npoly; j++) {
      L.src[j] = dst[j];
      L.dst[j] = dst[j];
      L.src2[j] = src[j];
      L.bsrc[j] =
          reinterpret_cast<const int32_t *>(v + (size_t)j * mp->src_len * kLimb);
      L.btab[j] = mp->d_ftab;
      L.bs[j] = mp->src_len;
      for (int y = 0; y < mp->dst_len; y++)
        jobs.push_back({mp->dst_prime[y], j | (j << 2) | (j << 4) | (j << 6), y,
                        y, mp->src2_limb[y], 1, mp->pad[y], mp->inv_prod[y],
                        2 * y * mp->src_len});
    }
    run_jobs(c, true, L, jobs, st);
  }
  return launch_status("moddown");
}

// ---------------------------------------------------------------------------
// Context: key switching
// ---------------------------------------------------------------------------

// MultKeyNoModDown (Context.cpp:609-712): ModUp(x), key inner product,
// accum[q] += P * add. Workspace: v (num_q) + modup (beta * tot). With
// `tensor` (b1, a1, b2, a2) the key-switched polynomial is a1 (*) a2 and
// the fold adds P * (b1 b2, b1 a2 + a1 b2) (HMult).
int op_mult_key_no_moddown(const Context &c, int level, uint32_t *acc_b,
                           uint32_t *acc_a, const uint32_t *x,
                           const uint32_t *add_b, const uint32_t *add_a,
                           const uint32_t *const *key_b,
                           const uint32_t *const *key_a, uint32_t *ws,
                           hipStream_t st, const uint32_t *const *tensor) {
  const LevelPlan &lp = c.levels[level];
  uint32_t *v = ws;
  uint32_t *modup = ws + (size_t)lp.num_q * kLimb;
  int rc = op_modup(c, level, modup, tensor ? tensor[1] : x, v, st, false,
                    tensor ? tensor[3] : nullptr);
  if (rc) return rc;
  KeyAccumArgs A;
  std::memset(&A, 0, sizeof(A));
  A.modup = modup;
  A.ngroups = 0;
  for (int g = 0; g < lp.beta; g++) {
    if (lp.groups[g].s1 <= lp.groups[g].s0) continue;
    A.group[A.ngroups++] = g;
    A.own0[g] = lp.groups[g].s0;
    A.own1[g] = lp.groups[g].s1;
    A.key_b[g] = key_b[g];
    A.key_a[g] = key_a[g];
  }
  A.x = x;
  A.add_b = add_b;
  A.add_a = add_a;
  if (tensor) {
    for (int i = 0; i < 4; i++) A.tensor[i] = tensor[i];
    A.use_tensor = 1;
  }
  A.p_prod = c.d_p_prod;
  A.out_b = acc_b;
  A.out_a = acc_a;
  EwNp np = make_np(c, lp.nm, lp.nt, c.alpha);
  launch_key_accum(c, np, A, st);
  return launch_status("key accumulate");
}

size_t mult_key_ws_limbs(const Context &c, int level) {
  const LevelPlan &lp = c.levels[level];
  return (size_t)lp.num_q + (size_t)lp.beta * (lp.num_q + c.alpha);
}

// Relinearize / RelinearizeRescale of (d0, d1, d2) (Context.cpp:561-565,
// :584-607, :714-737)
int op_relinearize(const Context &c, int level, uint32_t *out_b,
                   uint32_t *out_a, const uint32_t *d0, const uint32_t *d1,
                   const uint32_t *d2, const uint32_t *const *key_b,
                   const uint32_t *const *key_a, bool rescale, uint32_t *ws,
                   hipStream_t st) {
  const LevelPlan &lp = c.levels[level];
  const int tot = lp.num_q + c.alpha;
  uint32_t *acc_b = ws;
  uint32_t *acc_a = ws + (size_t)tot * kLimb;
  uint32_t *rest = ws + 2 * (size_t)tot * kLimb;
  int rc = op_mult_key_no_moddown(c, level, acc_b, acc_a, d2, d0, d1, key_b,
                                  key_a, rest, st);
  if (rc) return rc;
  uint32_t *dst[2] = {out_b, out_a};
  const uint32_t *src[2] = {acc_b, acc_a};
  return op_moddown(c, level, rescale ? 2 : 0, 2, dst, src, rest, st);
}

// HMult (Context.cpp:862-871): Tensor then Relinearize(Rescale), with the
// Tensor folded into the ModUp INTT load and the P * (d0, d1) fold.
int op_hmult(const Context &c, int level, uint32_t *out_b, uint32_t *out_a,
             const uint32_t *b1, const uint32_t *a1, const uint32_t *b2,
             const uint32_t *a2, const uint32_t *const *key_b,
             const uint32_t *const *key_a, bool rescale, hipStream_t st) {
  const LevelPlan &lp = c.levels[level];
  const int tot = lp.num_q + c.alpha;
  uint32_t *acc_b = c.ws;
  uint32_t *acc_a = c.ws + (size_t)tot * kLimb;
  uint32_t *rest = c.ws + 2 * (size_t)tot * kLimb;
  const uint32_t *tensor[4] = {b1, a1, b2, a2};
  int rc = op_mult_key_no_moddown(c, level, acc_b, acc_a, nullptr, nullptr,
                                  nullptr, key_b, key_a, rest, st, tensor);
  if (rc) return rc;
  uint32_t *dst[2] = {out_b, out_a};
  const uint32_t *src[2] = {acc_b, acc_a};
  return op_moddown(c, level, rescale ? 2 : 0, 2, dst, src, rest, st);
}

// MultKey on (b, a) (Context.cpp:584-607): key-switch a, accum.b += P b,
// ModDown both. Result into out (2 polys of num_q limbs).
int op_mult_key(const Context &c, int level, uint32_t *out_b, uint32_t *out_a,
                const uint32_t *b, const uint32_t *a,
                const uint32_t *const *key_b, const uint32_t *const *key_a,
                uint32_t *ws, hipStream_t st) {
  const LevelPlan &lp = c.levels[level];
  const int tot = lp.num_q + c.alpha;
  uint32_t *acc_b = ws;
  uint32_t *acc_a = ws + (size_t)tot * kLimb;
  uint32_t *rest = ws + 2 * (size_t)tot * kLimb;
  int rc = op_mult_key_no_moddown(c, level, acc_b, acc_a, a, b, nullptr, key_b,
                                  key_a, rest, st);
  if (rc) return rc;
  uint32_t *dst[2] = {out_b, out_a};
  const uint32_t *src[2] = {acc_b, acc_a};
  return op_moddown(c, level, 0, 2, dst, src, rest, st);
}

// HRot / HConj (Context.cpp:776-798): MultKey then Permute by galois factor
int op_hrot_galois(const Context &c, int level, uint32_t *out_b,
                   uint32_t *out_a, const uint32_t *b, const uint32_t *a,
                   const uint32_t *const *key_b, const uint32_t *const *key_a,
                   uint32_t galois, hipStream_t st) {
  const LevelPlan &lp = c.levels[level];
  const size_t qs = (size_t)lp.num_q * kLimb;
  uint32_t *tmp = c.ws;
  int rc = op_mult_key(c, level, tmp, tmp + qs, b, a, key_b, key_a,
                       tmp + 2 * qs, st);
  if (rc) return rc;
  EwNp np = make_np(c, lp.nm, lp.nt, 0);
  EwOperand x{{tmp, tmp + qs, nullptr}, 0};
  EwOperand o{{out_b, out_a, nullptr}, 0};
  launch_permute(c, np, x, o, galois, 2, st);
  return launch_status("permute");
}

}  // namespace ck
