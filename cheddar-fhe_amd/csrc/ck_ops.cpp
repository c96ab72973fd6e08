// Host orchestration of the key-switching path: ModUp, ModDown / Rescale /
// ModDownAndRescale, the key inner product and the evaluator compositions
// HMult, Relinearize, MultKey, HRot, HConj and hoisted rotations
// (reference: src/core/ModSwitch.cu:413-684, src/core/Context.cpp:316-871,
// src/extension/Hoist.cu:207-327, :578-691).
//
// Every launch plan is built on the host; temporaries come from a caller-
// provided workspace, so an operation issues no allocation, no host<->device
// copy and no synchronisation.
//
// Fusions relative to the reference's kernel sequence (results identical):
//   * BaseConv runs inside the forward NTT's column pass (ModUp and ModDown),
//   * HMult's Tensor product a1 (*) a2 is formed while the ModUp INTT loads,
//     and d0 = b1 (*) b2, d1 = b1 (*) a2 + a1 (*) b2 inside the key inner
//     product's P * (d0, d1) fold,
//   * ModUp's own-group limbs are read from the source by the inner product
//     instead of being copied,
//   * all polynomials of a ModDown share its launches.
#include "ck_ops.h"
#include "ck_i8.h"

#include <algorithm>
#include <cstring>

namespace ck {

namespace {

constexpr size_t kLimb = 65536;

int launch_status(const char *what) {
  return check_hip(hipGetLastError(), what);
}

uint32_t r_inv_mod(uint32_t q) { return h_invmod(h_to_mont(1, q), q); }

// slot of `ptr` among the first `n` entries of `arr` (appended if absent and
// there is room); -1 when all four slots hold other pointers
template <typename T>
int slot_of(T *arr, int &n, T ptr) {
  for (int i = 0; i < n; i++)
    if (arr[i] == ptr) return i;
  if (n == 4) return -1;
  arr[n] = ptr;
  return n++;
}

}  // namespace

// ---------------------------------------------------------------------------
// ImmediateSink
// ---------------------------------------------------------------------------

int ImmediateSink::pass(bool forward, int flags,
                        const std::vector<PassJob> &jobs) {
  NttLaunch L;
  int ns = 0, nd = 0, n2 = 0, nb = 0;
  auto reset = [&]() {
    std::memset(&L, 0, sizeof(L));
    L.flags = flags;
    ns = nd = n2 = nb = 0;
  };
  auto flush = [&]() {
    if (L.n == 0) return;
    if (forward)
      launch_ntt_fwd(c_, L, st_);
    else
      launch_intt(c_, L, st_);
    reset();
  };
  reset();
  for (const PassJob &j : jobs) {
    for (int attempt = 0;; attempt++) {
      int a = slot_of(L.src, ns, j.src);
      int b = slot_of(L.dst, nd, j.dst);
      int s2 = slot_of(L.src2, n2, j.src2);
      // BaseConv sources and table go together
      int bc = -1;
      for (int i = 0; i < nb; i++)
        if (L.bsrc[i] == j.bsrc && L.btab[i] == j.btab && L.bs[i] == j.bs)
          bc = i;
      if (bc < 0 && nb < 4) {
        L.bsrc[nb] = j.bsrc;
        L.btab[nb] = j.btab;
        L.bs[nb] = j.bs;
        bc = nb++;
      }
      if (a >= 0 && b >= 0 && s2 >= 0 && bc >= 0 && L.n < kMaxLaunchLimbs) {
        NttLimb &R = L.limb[L.n++];
        R.prime = (uint8_t)j.prime;
        R.sel = (uint8_t)(a | (b << 2) | (s2 << 4) | (bc << 6));
        R.src_limb = (uint16_t)j.src_limb;
        R.dst_limb = (uint16_t)j.dst_limb;
        R.src2_limb = (int16_t)(j.src2 ? j.src2_limb : -1);
        R.btab_off = (uint16_t)j.btab_off;
        R.c1 = j.c1;
        R.c2 = j.c2;
        R.c3 = j.c3;
        break;
      }
      if (attempt > 0) {
        set_error("transform job does not fit an empty launch");
        return kErrUnsupported;
      }
      flush();
    }
  }
  flush();
  return launch_status(forward ? "ntt" : "intt");
}

int ImmediateSink::copy(uint32_t *dst, const uint32_t *src, size_t words) {
  return launch_copy(c_, dst, src, words * sizeof(uint32_t), st_);
}

int ImmediateSink::bconv(const std::vector<BconvInst> &insts) {
  if (c_.bconv_mfma == 2) {  // int8 matrix cores
    for (size_t i = 0; i < insts.size(); i += kMaxBconvI8Inst) {
      BconvI8Launch L;
      std::memset(&L, 0, sizeof(L));
      L.n = (int)std::min<size_t>(kMaxBconvI8Inst, insts.size() - i);
      for (int k = 0; k < L.n; k++)
        if (!to_i8_inst(c_, insts[i + k], &L.inst[k])) return kErrHip;
      launch_bconv_i8(c_, L, st_);
    }
    return launch_status("bconv");
  }
  for (size_t i = 0; i < insts.size(); i += kMaxBconvInst) {
    BconvLaunch L;
    std::memset(&L, 0, sizeof(L));
    L.n = (int)std::min<size_t>(kMaxBconvInst, insts.size() - i);
    for (int k = 0; k < L.n; k++) L.inst[k] = insts[i + k];
    launch_bconv_mfma(c_, L, st_);
  }
  return launch_status("bconv");
}

int ImmediateSink::key_rows(const EwNp &np, const KeyAccumArgs &A) {
  launch_key_rows(c_, np, A, st_);
  return launch_status("key rows");
}

int ImmediateSink::key_accum(const EwNp &np, const KeyAccumArgs &A) {
  launch_key_accum(c_, np, A, st_);
  return launch_status("key accumulate");
}

int ImmediateSink::bs_fused(const EwNp &np, const BsFusedArgs &A) {
  launch_bs_fused(c_, np, A, st_);
  return launch_status("hoisted key accumulate");
}

int ImmediateSink::permute(const EwNp &np, const EwOperand &x,
                           const EwOperand &out, uint32_t g, int npoly) {
  launch_permute(c_, np, x, out, g, npoly, st_);
  return launch_status("permute");
}

// ---------------------------------------------------------------------------
// NTTHandler
// ---------------------------------------------------------------------------

int op_ntt(const Context &c, uint32_t *dst, const uint32_t *src, int nm,
           int nt, int na, int src_gap, bool mont_conv, Sink &s) {
  const int num_q = nm + nt;
  std::vector<PassJob> jobs;
  for (int t = 0; t < num_q + na; t++) {
    const int p = global_prime(c, nm, nt, t);
    PassJob j;
    j.src = src;
    j.src_limb = t < num_q ? t : t + src_gap;
    j.dst = dst;
    j.dst_limb = t;
    j.prime = p;
    // NTT(..., montgomery_conversion): x * R^2 * R^-1 = x * R
    j.c1 = mont_conv ? h_to_mont(1, c.primes[p]) : 1u;
    jobs.push_back(j);
  }
  return s.pass(true, mont_conv ? kNttPreMul : 0, jobs);
}

// INTT followed by x * const * R^-1 with per-limb constants given in the
// reference's convention (the value it passes to MultConst).
int op_intt_const(const Context &c, uint32_t *dst, const uint32_t *src,
                  int nm, int nt, int na, int src_gap,
                  const uint32_t *consts_mont, bool centred, Sink &s) {
  const int num_q = nm + nt;
  std::vector<PassJob> jobs;
  for (int t = 0; t < num_q + na; t++) {
    const int p = global_prime(c, nm, nt, t);
    const uint32_t q = c.primes[p];
    PassJob j;
    j.src = src;
    j.src_limb = t < num_q ? t : t + src_gap;
    j.dst = dst;
    j.dst_limb = t;
    j.prime = p;
    j.c1 = h_mulmod(consts_mont[t] % q, r_inv_mod(q), q);
    jobs.push_back(j);
  }
  return s.pass(false, centred ? kInttCentred : 0, jobs);
}

int op_intt(const Context &c, uint32_t *dst, const uint32_t *src, int nm,
            int nt, int na, int src_gap, bool mont_conv, Sink &s) {
  // NTT.cu:673-741: N^-1 (leave Montgomery form) or N^-1 * R (keep it)
  std::vector<uint32_t> cs;
  for (int t = 0; t < nm + nt + na; t++) {
    const uint32_t q = c.primes[global_prime(c, nm, nt, t)];
    const uint32_t ninv = h_invmod((uint32_t)c.n % q, q);
    cs.push_back(mont_conv ? ninv : h_to_mont(ninv, q));
  }
  return op_intt_const(c, dst, src, nm, nt, na, src_gap, cs.data(), false, s);
}

// ---------------------------------------------------------------------------
// ModSwitchHandler
// ---------------------------------------------------------------------------

// ModUp (ModSwitch.cu:429-509). dst holds beta blocks of (num_q + alpha)
// limbs; v is a workspace of num_q limbs. When prod2 is given the source is
// src (*) prod2 (HMult's a1 (*) a2). copy_own = false leaves each group's own
// limbs unwritten (the fused inner product reads them from the source).
int op_modup(const Context &c, int level, uint32_t *dst, const uint32_t *src,
             uint32_t *v, Sink &s, bool copy_own, const uint32_t *prod2,
             bool cols_only) {
  const LevelPlan &lp = level_plan(c, level);
  const int tot = lp.num_q + level_naux(c, level);
  // 1. INTT of all q limbs with the per-group constant, centred output
  {
    std::vector<PassJob> jobs;
    for (int t = 0; t < lp.num_q; t++) {
      PassJob j;
      j.src = src;
      j.src_limb = t;
      j.dst = v;
      j.dst_limb = t;
      j.src2 = prod2;
      j.src2_limb = t;
      j.prime = lp.q_prime[t];
      j.c1 = lp.modup_c1[t];
      jobs.push_back(j);
    }
    int rc = s.pass(false, kInttCentred | (prod2 ? kInttLoadProduct : 0), jobs);
    if (rc) return rc;
  }
  // 2. BaseConv of each group, on the matrix cores or fused into the
  // forward NTT of its outputs
  const bool mfma = s.mfma_bconv();
  std::vector<PassJob> jobs;
  std::vector<BconvInst> insts;
  for (int g = 0; g < lp.beta; g++) {
    const ModUpGroup &gr = lp.groups[g];
    if (gr.s1 <= gr.s0) continue;
    const int sl = gr.s1 - gr.s0;
    const int32_t *bsrc = reinterpret_cast<const int32_t *>(v + (size_t)gr.s0 * kLimb);
    if (mfma) {
      BconvInst I;
      std::memset(&I, 0, sizeof(I));
      I.src = bsrc;
      I.dst = dst;
      I.tab = gr.d_ftab;
      I.S = sl;
      I.D = (int)gr.dst_limb.size();
      for (int m = 0; m < I.D; m++) {
        I.dst_limb[m] = (uint16_t)(g * tot + gr.dst_limb[m]);
        I.dst_prime[m] = (uint8_t)gr.dst_prime[m];
      }
      insts.push_back(I);
    }
    for (size_t m = 0; m < gr.dst_limb.size(); m++) {
      PassJob j;
      j.src = dst;
      j.dst = dst;
      j.src_limb = j.dst_limb = g * tot + gr.dst_limb[m];
      if (!mfma) {
        j.bsrc = bsrc;
        j.btab = gr.d_ftab;
        j.bs = sl;
        j.btab_off = 2 * (int)m * sl;
      }
      j.prime = gr.dst_prime[m];
      j.c1 = 1;
      jobs.push_back(j);
    }
  }
  int rc = mfma ? s.bconv(insts) : kOk;
  if (!rc)
    rc = s.pass(true, (mfma ? 0 : kNttBconvInput) | (cols_only ? kNttColsOnly : 0), jobs);
  if (rc) return rc;
  if (copy_own) {
    for (int g = 0; g < lp.beta; g++) {
      const ModUpGroup &gr = lp.groups[g];
      if (gr.s1 <= gr.s0) continue;
      rc = s.copy(dst + ((size_t)g * tot + gr.s0) * kLimb,
                  src + (size_t)gr.s0 * kLimb, (size_t)(gr.s1 - gr.s0) * kLimb);
      if (rc) return rc;
    }
  }
  return kOk;
}

const ModDownPlan *moddown_plan(const Context &c, int level, int kind) {
  if (!has_switch_level(c, level) || (level < 0 && kind != 0)) return nullptr;
  const LevelPlan &lp = level_plan(c, level);
  const ModDownPlan *mp = kind == 0   ? &lp.moddown
                          : kind == 1 ? &lp.rescale
                                      : &lp.moddown_rescale;
  return mp->valid ? mp : nullptr;
}

// ModDownWorker (ModSwitch.cu:566-684) for npoly polynomials at once; v is
// a workspace of npoly * src_len limbs.
int op_moddown(const Context &c, int level, int kind, int npoly,
               uint32_t *const *dst, const uint32_t *const *src, uint32_t *v,
               Sink &s) {
  const ModDownPlan *mp = moddown_plan(c, level, kind);
  if (!mp) {
    set_error("no ModDown plan for this level and kind");
    return kErrInvalidArg;
  }
  // 1. INTT of the removed limbs, x const1, centred
  {
    std::vector<PassJob> jobs;
    for (int k = 0; k < npoly; k++)
      for (int i = 0; i < mp->src_len; i++) {
        PassJob j;
        j.src = src[k];
        j.src_limb = mp->src_limb[i];
        j.dst = v + (size_t)k * mp->src_len * kLimb;
        j.dst_limb = i;
        j.prime = mp->src_prime[i];
        j.c1 = mp->c1[i];
        jobs.push_back(j);
      }
    int rc = s.pass(false, kInttCentred, jobs);
    if (rc) return rc;
  }
  // 2. BaseConv (matrix cores, or fused into the forward NTT), 3. epilogue
  // in the row pass
  const bool mfma = s.mfma_bconv();
  std::vector<PassJob> jobs;
  std::vector<BconvInst> insts;
  for (int k = 0; k < npoly; k++) {
    const int32_t *bsrc =
        reinterpret_cast<const int32_t *>(v + (size_t)k * mp->src_len * kLimb);
    if (mfma) {
      BconvInst I;
      std::memset(&I, 0, sizeof(I));
      I.src = bsrc;
      I.dst = dst[k];
      I.tab = mp->d_ftab;
      I.S = mp->src_len;
      I.D = mp->dst_len;
      for (int y = 0; y < I.D; y++) {
        I.dst_limb[y] = (uint16_t)y;
        I.dst_prime[y] = (uint8_t)mp->dst_prime[y];
      }
      insts.push_back(I);
    }
    for (int y = 0; y < mp->dst_len; y++) {
      PassJob j;
      j.src = dst[k];
      j.dst = dst[k];
      j.src_limb = j.dst_limb = y;
      j.src2 = src[k];
      j.src2_limb = mp->src2_limb[y];
      if (!mfma) {
        j.bsrc = bsrc;
        j.btab = mp->d_ftab;
        j.bs = mp->src_len;
        j.btab_off = 2 * y * mp->src_len;
      }
      j.prime = mp->dst_prime[y];
      j.c1 = 1;
      j.c2 = mp->pad[y];
      j.c3 = mp->inv_prod[y];
      jobs.push_back(j);
    }
  }
  if (mfma) {
    int rc = s.bconv(insts);
    if (rc) return rc;
  }
  return s.pass(true,
                (mfma ? 0 : kNttBconvInput) | kNttModDownEpilogue |
                    (mp->use_pad ? kNttEpiloguePad : 0),
                jobs);
}

// ---------------------------------------------------------------------------
// Context: key switching
// ---------------------------------------------------------------------------

namespace {

// the ModUp groups that carry limbs at this level, in KeyAccumArgs form
void fill_groups(const LevelPlan &lp, const uint32_t *const *key_b,
                 const uint32_t *const *key_a, KeyAccumArgs &A) {
  A.ngroups = 0;
  for (int g = 0; g < lp.beta; g++) {
    if (lp.groups[g].s1 <= lp.groups[g].s0) continue;
    A.group[A.ngroups++] = g;
    A.own0[g] = lp.groups[g].s0;
    A.own1[g] = lp.groups[g].s1;
    A.key_b[g] = key_b[g];
    A.key_a[g] = key_a[g];
  }
}

}  // namespace

size_t mult_key_ws_limbs(const Context &c, int level) {
  const LevelPlan &lp = level_plan(c, level);
  return (size_t)lp.num_q + (size_t)lp.beta * (lp.num_q + level_naux(c, level));
}

// MultKeyNoModDown (Context.cpp:609-712): ModUp(x), key inner product,
// accum[q] += P * add. Workspace: v (num_q) + modup (beta * tot). With
// `tensor` (b1, a1, b2, a2) the key-switched polynomial is a1 (*) a2 and
// the fold adds P * (b1 b2, b1 a2 + a1 b2) (HMult).
int op_mult_key_no_moddown(const Context &c, int level, uint32_t *acc_b,
                           uint32_t *acc_a, const uint32_t *x,
                           const uint32_t *add_b, const uint32_t *add_a,
                           const uint32_t *const *key_b,
                           const uint32_t *const *key_a, uint32_t *ws,
                           Sink &s, const uint32_t *const *tensor) {
  const LevelPlan &lp = level_plan(c, level);
  const int naux = level_naux(c, level);
  uint32_t *v = ws;
  uint32_t *modup = ws + (size_t)lp.num_q * kLimb;
  const bool fused = s.fused_key_rows(c);
  int rc = op_modup(c, level, modup, tensor ? tensor[1] : x, v, s, false,
                    tensor ? tensor[3] : nullptr, fused);
  if (rc) return rc;
  KeyAccumArgs A;
  std::memset(&A, 0, sizeof(A));
  A.modup = modup;
  fill_groups(lp, key_b, key_a, A);
  A.x = x;
  A.add_b = add_b;
  A.add_a = add_a;
  if (tensor) {
    for (int i = 0; i < 4; i++) A.tensor[i] = tensor[i];
    A.use_tensor = 1;
  }
  // level -1: the key block holds the short base's q limbs then its P'
  // limbs (the key's own NP, prime offset 0: Context.cpp:619-646); other
  // levels: full-size blocks indexed by global prime
  if (level < 0) {
    A.key_q0 = 0;
    A.key_aux0 = lp.num_q;
    A.p_prod = c.d_p_prod_short;
  } else {
    A.key_q0 = c.max_ter - lp.nt;
    A.key_aux0 = c.L;
    A.p_prod = c.d_p_prod;
  }
  A.out_b = acc_b;
  A.out_a = acc_a;
  const EwNp np = make_np(c, lp.nm, lp.nt, naux);
  return fused ? s.key_rows(np, A) : s.key_accum(np, A);
}

size_t relin_ws_limbs(const Context &c, int level) {
  const LevelPlan &lp = level_plan(c, level);
  const size_t tot = (size_t)lp.num_q + level_naux(c, level);
  // accumulators + max(ModUp workspace, ModDown workspace of two polys)
  return 2 * tot + std::max(mult_key_ws_limbs(c, level), 2 * (size_t)tot);
}

// Relinearize / RelinearizeRescale of (d0, d1, d2) (Context.cpp:561-565,
// :584-607, :714-737)
int op_relinearize(const Context &c, int level, uint32_t *out_b,
                   uint32_t *out_a, const uint32_t *d0, const uint32_t *d1,
                   const uint32_t *d2, const uint32_t *const *key_b,
                   const uint32_t *const *key_a, bool rescale, uint32_t *ws,
                   Sink &s) {
  const LevelPlan &lp = c.levels[level];
  const int tot = lp.num_q + c.alpha;
  uint32_t *acc_b = ws;
  uint32_t *acc_a = ws + (size_t)tot * kLimb;
  uint32_t *rest = ws + 2 * (size_t)tot * kLimb;
  int rc = op_mult_key_no_moddown(c, level, acc_b, acc_a, d2, d0, d1, key_b,
                                  key_a, rest, s);
  if (rc) return rc;
  uint32_t *dst[2] = {out_b, out_a};
  const uint32_t *src[2] = {acc_b, acc_a};
  return op_moddown(c, level, rescale ? 2 : 0, 2, dst, src, rest, s);
}

// HMult (Context.cpp:862-871): Tensor then Relinearize(Rescale), with the
// Tensor folded into the ModUp INTT load and the P * (d0, d1) fold.
int op_hmult(const Context &c, int level, uint32_t *out_b, uint32_t *out_a,
             const uint32_t *b1, const uint32_t *a1, const uint32_t *b2,
             const uint32_t *a2, const uint32_t *const *key_b,
             const uint32_t *const *key_a, bool rescale, uint32_t *ws,
             Sink &s) {
  const LevelPlan &lp = c.levels[level];
  const int tot = lp.num_q + c.alpha;
  uint32_t *acc_b = ws;
  uint32_t *acc_a = ws + (size_t)tot * kLimb;
  uint32_t *rest = ws + 2 * (size_t)tot * kLimb;
  const uint32_t *tensor[4] = {b1, a1, b2, a2};
  int rc = op_mult_key_no_moddown(c, level, acc_b, acc_a, nullptr, nullptr,
                                  nullptr, key_b, key_a, rest, s, tensor);
  if (rc) return rc;
  uint32_t *dst[2] = {out_b, out_a};
  const uint32_t *src[2] = {acc_b, acc_a};
  return op_moddown(c, level, rescale ? 2 : 0, 2, dst, src, rest, s);
}

// MultKey on (b, a) (Context.cpp:584-607): key-switch a, accum.b += P b,
// ModDown both. Result into out (2 polys of num_q limbs).
int op_mult_key(const Context &c, int level, uint32_t *out_b, uint32_t *out_a,
                const uint32_t *b, const uint32_t *a,
                const uint32_t *const *key_b, const uint32_t *const *key_a,
                uint32_t *ws, Sink &s) {
  const LevelPlan &lp = level_plan(c, level);
  const int tot = lp.num_q + level_naux(c, level);
  uint32_t *acc_b = ws;
  uint32_t *acc_a = ws + (size_t)tot * kLimb;
  uint32_t *rest = ws + 2 * (size_t)tot * kLimb;
  int rc = op_mult_key_no_moddown(c, level, acc_b, acc_a, a, b, nullptr, key_b,
                                  key_a, rest, s);
  if (rc) return rc;
  uint32_t *dst[2] = {out_b, out_a};
  const uint32_t *src[2] = {acc_b, acc_a};
  return op_moddown(c, level, 0, 2, dst, src, rest, s);
}

// HRot / HConj (Context.cpp:776-798): MultKey then Permute by galois factor
int op_hrot_galois(const Context &c, int level, uint32_t *out_b,
                   uint32_t *out_a, const uint32_t *b, const uint32_t *a,
                   const uint32_t *const *key_b, const uint32_t *const *key_a,
                   uint32_t galois, uint32_t *ws, Sink &s) {
  const LevelPlan &lp = c.levels[level];
  const size_t qs = (size_t)lp.num_q * kLimb;
  uint32_t *tmp = ws;
  int rc = op_mult_key(c, level, tmp, tmp + qs, b, a, key_b, key_a,
                       tmp + 2 * qs, s);
  if (rc) return rc;
  EwNp np = make_np(c, lp.nm, lp.nt, 0);
  EwOperand x{{tmp, tmp + qs, nullptr}, 0};
  EwOperand o{{out_b, out_a, nullptr}, 0};
  return s.permute(np, x, o, galois, 2);
}

// ---------------------------------------------------------------------------
// HoistHandler: hoisted rotations
// ---------------------------------------------------------------------------

size_t hoisted_ws_limbs(const Context &c, int level, int chunk_rot,
                        bool moddown) {
  const LevelPlan &lp = c.levels[level];
  const size_t tot = (size_t)lp.num_q + c.alpha;
  size_t need = mult_key_ws_limbs(c, level);
  if (moddown) need += 2 * (size_t)chunk_rot * (tot + c.alpha);
  return need;
}

// EvaluateBabyStep (Hoist.cu:578-691) with the fused key product
// (BSFusedKeyMult :207-327), followed, when `moddown`, by the ModDown of
// every baby-step output (the MultKey tail of HRot): the results equal
// HRot(ct, rots[i]) bit for bit, because ModDown commutes with the
// automorphism and ModDown(P b + s) = b + ModDown(s). The workspace holds
// the accumulators of chunk_rot rotations; each chunk's outputs share one
// ModDown.
int op_hoisted_rotations(const Context &c, int level, int n_rot,
                         const int *rots, uint32_t *const *out_b,
                         uint32_t *const *out_a, const uint32_t *b,
                         const uint32_t *a, const uint32_t *const *key_b,
                         const uint32_t *const *key_a, bool moddown,
                         int chunk_rot, uint32_t *ws, Sink &s) {
  const LevelPlan &lp = c.levels[level];
  const int tot = lp.num_q + c.alpha;
  uint32_t *v = ws;
  uint32_t *modup = ws + (size_t)lp.num_q * kLimb;
  uint32_t *acc = modup + (size_t)lp.beta * tot * kLimb;
  int rc = op_modup(c, level, modup, a, v, s, false);
  if (rc) return rc;
  BsFusedArgs A;
  std::memset(&A, 0, sizeof(A));
  A.modup = modup;
  A.x = a;
  A.b = b;
  A.p_prod = c.d_p_prod;
  for (int g = 0; g < lp.beta; g++) {
    if (lp.groups[g].s1 <= lp.groups[g].s0) continue;
    if (A.ngroups == kMaxBsGroups) {
      set_error("hoisted rotations support at most 8 key groups");
      return kErrUnsupported;
    }
    A.group[A.ngroups] = g;
    A.own0[A.ngroups] = lp.groups[g].s0;
    A.own1[A.ngroups] = lp.groups[g].s1;
    A.ngroups++;
  }
  const EwNp np = make_np(c, lp.nm, lp.nt, c.alpha);
  if (chunk_rot < 1) chunk_rot = 1;
  for (int c0 = 0; c0 < n_rot; c0 += chunk_rot) {
    const int nc = std::min(chunk_rot, n_rot - c0);
    std::vector<uint32_t *> md_dst;
    std::vector<const uint32_t *> md_src;
    for (int r0 = c0; r0 < c0 + nc; r0 += kMaxBsRot) {
      const int nr = std::min(kMaxBsRot, c0 + nc - r0);
      A.nrot = nr;
      for (int r = 0; r < nr; r++) {
        const int i = r0 + r;
        // scatter by g^-1 = 5^(N/2 - rot) (Hoist.cu:300-302)
        A.ginv[r] = c.galois[c.n / 2 - rots[i]];
        for (int gi = 0; gi < A.ngroups; gi++) {
          const int g = A.group[gi];
          A.key_b[r][gi] = key_b[(size_t)i * c.dnum + g];
          A.key_a[r][gi] = key_a[(size_t)i * c.dnum + g];
        }
        if (moddown) {
          A.out_b[r] = acc + (size_t)(2 * (i - c0)) * tot * kLimb;
          A.out_a[r] = acc + (size_t)(2 * (i - c0) + 1) * tot * kLimb;
          md_src.push_back(A.out_b[r]);
          md_src.push_back(A.out_a[r]);
          md_dst.push_back(out_b[i]);
          md_dst.push_back(out_a[i]);
        } else {
          A.out_b[r] = out_b[i];
          A.out_a[r] = out_a[i];
        }
      }
      rc = s.bs_fused(np, A);
      if (rc) return rc;
    }
    if (moddown) {
      uint32_t *mv = acc + (size_t)(2 * nc) * tot * kLimb;
      rc = op_moddown(c, level, 0, 2 * nc, md_dst.data(), md_src.data(), mv, s);
      if (rc) return rc;
    }
  }
  return kOk;
}

}  // namespace ck
