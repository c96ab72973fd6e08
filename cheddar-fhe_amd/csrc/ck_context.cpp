// Context construction: prime tables, twiddles and the per-level ModUp /
// ModDown / Rescale / ModDownAndRescale constant plans, computed once on the
// host and uploaded (reference: Parameter ctor src/core/Parameter.cu:20-193,
// NTTHandler::PopulateTwiddleFactors src/core/NTT.cu:1233-1299,
// ModSwitchHandler ctor + PopulateModSwitchConstants +
// PopulateModDownEpilogueConstants src/core/ModSwitch.cu:175-411, Context
// ctor P table src/core/Context.cpp:85-98).
//
// The kernels compute transforms in exact FP64 on plain representatives, so
// every constant that the reference applies through a Montgomery product
// (x * c * 2^-32) is stored here pre-multiplied by 2^-32 mod q ("plain
// form"); the outputs are the same canonical residues.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "ck_internal.h"
#include "ck_ops.h"

namespace ck {

namespace {
thread_local std::string g_err;
}

void set_error(const std::string &msg) { g_err = msg; }
const char *last_error() { return g_err.c_str(); }

int check_hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return kOk;
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? kErrOutOfMemory : kErrHip;
}

// ---------------------------------------------------------------------------
// Host number theory
// ---------------------------------------------------------------------------
uint32_t h_mulmod(uint32_t a, uint32_t b, uint32_t q) {
  return (uint32_t)(((uint64_t)a * b) % q);
}

uint32_t h_powmod(uint32_t a, uint64_t e, uint32_t q) {
  uint64_t r = 1 % q, b = a % q;
  while (e) {
    if (e & 1) r = r * b % q;
    b = b * b % q;
    e >>= 1;
  }
  return (uint32_t)r;
}

// q prime: a^(q-2)
uint32_t h_invmod(uint32_t a, uint32_t q) { return h_powmod(a, q - 2, q); }

uint32_t h_to_mont(uint32_t a, uint32_t q) {
  return (uint32_t)(((uint64_t)(a % q) << 32) % q);
}

namespace {

int32_t inv_mod_2_32(uint32_t q) {
  uint32_t x = q;
  for (int i = 0; i < 5; i++) x *= 2u - q * x;
  return (int32_t)x;
}

// smallest generator of Z_q^* (the reference's FindGenerator picks the
// smallest; psi depends on that choice)
uint32_t smallest_generator(uint32_t q) {
  uint32_t phi = q - 1, rest = phi;
  std::vector<uint32_t> f;
  for (uint32_t d = 2; (uint64_t)d * d <= rest; d++) {
    if (rest % d == 0) {
      f.push_back(d);
      while (rest % d == 0) rest /= d;
    }
  }
  if (rest > 1) f.push_back(rest);
  for (uint32_t g = 2; g < q; g++) {
    bool ok = true;
    for (uint32_t d : f)
      if (h_powmod(g, phi / d, q) == 1) {
        ok = false;
        break;
      }
    if (ok) return g;
  }
  return 0;
}

bool is_prime_u32(uint32_t n) {
  if (n < 2) return false;
  for (uint32_t d = 2; (uint64_t)d * d <= n; d++)
    if (n % d == 0) return false;
  return true;
}

uint32_t brv(uint32_t x, int bits) {
  return __builtin_bitreverse32(x) >> (32 - bits);
}

int32_t centred(uint32_t a, uint32_t q) {
  return a > (q >> 1) ? (int32_t)a - (int32_t)q : (int32_t)a;
}

// r^-1 = (2^32)^-1 mod q
uint32_t r_inv(uint32_t q) { return h_invmod(h_to_mont(1, q), q); }

template <typename T>
int upload(Context &c, T **dptr, const std::vector<T> &h) {
  size_t bytes = std::max<size_t>(1, h.size()) * sizeof(T);
  int rc = check_hip(hipMalloc((void **)dptr, bytes), "hipMalloc(table)");
  if (rc) return rc;
  c.allocations.push_back(*dptr);
  if (!h.empty())
    rc = check_hip(hipMemcpy(*dptr, h.data(), h.size() * sizeof(T),
                             hipMemcpyHostToDevice),
                   "hipMemcpy(table)");
  return rc;
}

}  // namespace

// PopulateModSwitchConstants (ModSwitch.cu:320-366), kFuseModDownEpilogue:
// c1[i] = ((prod_{j!=i} s_j) * N)^-1 * prod(restore) mod s_i, returned in
// plain form (times R^-1 because the reference applies it with a Montgomery
// product); table[m][i] = centred((prod_{j!=i} s_j) * R^2 mod t_m).
//
// ftab holds the same products for the exact FP64 BaseConv fused into the
// forward NTT: T'[m][i] = (prod_{j!=i} s_j) * R mod t_m (one Montgomery
// factor: the fused form adds no R^-1) split as hi * 2^16 + lo.
void bconv_constants(const Context &c, const std::vector<int> &src,
                     const std::vector<int> &dst,
                     const std::vector<int> &restore,
                     std::vector<uint32_t> &c1, std::vector<int32_t> &table,
                     std::vector<double> &ftab) {
  const size_t s = src.size(), d = dst.size();
  c1.assign(s, 0);
  table.assign(s * d, 0);
  ftab.assign(2 * s * d, 0.0);
  for (size_t i = 0; i < s; i++) {
    uint32_t mod = c.primes[src[i]];
    uint32_t v = 1;
    for (size_t j = 0; j < s; j++)
      if (j != i) v = h_mulmod(v, c.primes[src[j]] % mod, mod);
    v = h_mulmod(v, (uint32_t)c.n % mod, mod);
    v = h_invmod(v, mod);
    for (int r : restore) v = h_mulmod(v, c.primes[r] % mod, mod);
    c1[i] = h_mulmod(v, r_inv(mod), mod);
  }
  for (size_t m = 0; m < d; m++) {
    uint32_t mod = c.primes[dst[m]];
    for (size_t i = 0; i < s; i++) {
      uint32_t v = 1;
      for (size_t j = 0; j < s; j++)
        if (j != i) v = h_mulmod(v, c.primes[src[j]] % mod, mod);
      const uint32_t v1 = h_to_mont(v, mod);
      ftab[2 * (m * s + i)] = (double)(v1 >> 16);
      ftab[2 * (m * s + i) + 1] = (double)(v1 & 0xFFFFu);
      table[m * s + i] = centred(h_to_mont(v1, mod), mod);
    }
  }
}

namespace {

// One ModDown flavour from (nm, nt) q primes with naux auxiliary primes to
// (dm, dt) (ModSwitch.cu:233-303 for the prime ranges, :566-684 for which
// limbs are transformed, NTT.cu:456-601 for the epilogue indexing). kind 0
// (ModDown) has dm = nm, dt = nt.
int build_moddown(Context &c, int nm, int nt, int dm, int dt, int naux, int kind,
                  ModDownPlan &mp) {
  const int num_q = nm + nt;
  const int next_num_q = dm + dt;
  const int num_src_aux = (kind == 1) ? 0 : naux;
  std::vector<int> dst_prime(next_num_q);
  for (int y = 0; y < next_num_q; y++) dst_prime[y] = c.max_ter - dt + y;

  std::vector<int> src_limb, src_prime, restore;
  int restore_start = 0, restore_end = 0;
  if (kind != 0) {
    const int main_diff = nm - dm, ter_diff = nt - dt;
    if (ter_diff > 0) {
      if (main_diff > 0) return kOk;  // not a valid rescale step
      for (int t = 0; t < ter_diff; t++) src_limb.push_back(t);
      restore_start = next_num_q + main_diff;
      restore_end = next_num_q;
    } else if (main_diff > 0) {
      for (int t = num_q - main_diff; t < num_q; t++) src_limb.push_back(t);
      restore_start = 0;
      restore_end = -ter_diff;
    } else {
      return kOk;
    }
  }
  for (int a = 0; a < num_src_aux; a++) src_limb.push_back(num_q + a);
  for (int t : src_limb)
    src_prime.push_back(t < num_q ? c.max_ter - nt + t : c.L + (t - num_q));
  for (int j = restore_start; j < restore_end; j++)
    restore.push_back(dst_prime[j]);

  std::vector<int32_t> table;
  std::vector<double> ftab;
  bconv_constants(c, src_prime, dst_prime, restore, mp.c1, table, ftab);
  mp.src_len = (int)src_limb.size();
  mp.dst_len = next_num_q;
  mp.src_limb = src_limb;
  mp.src_prime = src_prime;
  mp.dst_prime = dst_prime;
  if (mp.src_len > kMaxBconvSrc || mp.dst_len > kMaxBconvDst ||
      mp.dst_len > kMaxLaunchLimbs) {
    set_error("moddown plan exceeds kernel limits");
    return kErrUnsupported;
  }
  // epilogue: src2 rows and constants
  const int im = std::min(nm, dm), it = std::min(nt, dt);
  const int src2_start = dt - it, src2_end = src2_start + im + it;
  const int src2_offset = std::max(0, nt - dt);
  mp.src2_limb.assign(next_num_q, -1);
  mp.pad.assign(next_num_q, 1);
  mp.inv_prod.assign(next_num_q, 1);
  for (int y = 0; y < next_num_q; y++) {
    uint32_t mod = c.primes[dst_prime[y]];
    if (y >= src2_start && y < src2_end)
      mp.src2_limb[y] = src2_offset + (y - src2_start);
    uint32_t pv = 1;
    for (int r : restore) pv = h_mulmod(pv, c.primes[r] % mod, mod);
    mp.pad[y] = pv;
    uint32_t ip = 1;
    for (int sp : src_prime) ip = h_mulmod(ip, c.primes[sp] % mod, mod);
    mp.inv_prod[y] = h_invmod(ip, mod);
  }
  mp.use_pad = !restore.empty();
  int rc = upload(c, &mp.d_ftab, ftab);
  if (rc) return rc;
  mp.valid = true;
  return kOk;
}

// ModUp groups of naux q limbs each after `pad` padding limbs (the
// ModSwitchHandler constructor, ModSwitch.cu:175-232), for the q primes
// (nm, nt) and the first naux auxiliary primes.
int build_modup(Context &c, LevelPlan &lp, int nm, int nt, int naux, int pad) {
  lp.nm = nm;
  lp.nt = nt;
  lp.num_q = nm + nt;
  lp.pad = pad;
  const int padded = lp.num_q + pad;
  lp.beta = (padded + naux - 1) / naux;
  lp.q_prime.resize(lp.num_q);
  for (int t = 0; t < lp.num_q; t++) lp.q_prime[t] = c.max_ter - nt + t;
  lp.modup_c1.assign(lp.num_q, 0);
  lp.groups.assign(lp.beta, ModUpGroup());
  for (int g = 0; g < lp.beta; g++) {
    int s0 = g * naux, s1 = std::min((g + 1) * naux, padded);
    ModUpGroup &gr = lp.groups[g];
    if (s1 <= pad) {
      gr.s0 = gr.s1 = 0;
      continue;
    }
    s0 = std::max(0, s0 - pad);
    s1 -= pad;
    gr.s0 = s0;
    gr.s1 = s1;
    std::vector<int> src, dst;
    for (int t = s0; t < s1; t++) src.push_back(lp.q_prime[t]);
    for (int t = 0; t < lp.num_q + naux; t++) {
      if (t >= s0 && t < s1) continue;
      gr.dst_limb.push_back(t);
      dst.push_back(t < lp.num_q ? lp.q_prime[t] : c.L + (t - lp.num_q));
    }
    gr.dst_prime = dst;
    if ((int)src.size() > kMaxBconvSrc || (int)dst.size() > kMaxBconvDst) {
      set_error("modup group exceeds kernel limits");
      return kErrUnsupported;
    }
    std::vector<uint32_t> c1;
    std::vector<int32_t> table;
    std::vector<double> ftab;
    bconv_constants(c, src, dst, {}, c1, table, ftab);
    for (int t = s0; t < s1; t++) lp.modup_c1[t] = c1[t - s0];
    int rc = upload(c, &gr.d_ftab, ftab);
    if (rc) return rc;
  }
  return kOk;
}

int build_level(Context &c, int level) {
  LevelPlan &lp = c.levels[level];
  const int nm = c.lv_main[level], nt = c.lv_ter[level];
  // beta = ceil((nm + max_ter) / alpha): the q limbs padded at the front to
  // max_ter terminal limbs (ModSwitch.cu:181-185, :200-203)
  int rc = build_modup(c, lp, nm, nt, c.alpha, c.max_ter - nt);
  if (!rc) rc = build_moddown(c, nm, nt, nm, nt, c.alpha, 0, lp.moddown);
  if (level == 0) return rc;  // no rescale below level 0
  const int dm = c.lv_main[level - 1], dt = c.lv_ter[level - 1];
  if (!rc) rc = build_moddown(c, nm, nt, dm, dt, c.alpha, 1, lp.rescale);
  if (!rc) rc = build_moddown(c, nm, nt, dm, dt, c.alpha, 2, lp.moddown_rescale);
  return rc;
}

// CompareNPPair (Parameter.cu:10-15)
bool np_less(int am, int at, int bm, int bt) {
  return (am + at < bm + bt) || (am + at == bm + bt && am < bm);
}

}  // namespace

int global_prime(const Context &c, int num_main, int num_ter, int limb) {
  int num_q = num_main + num_ter;
  return limb < num_q ? c.max_ter - num_ter + limb : c.L + (limb - num_q);
}

// The level -1 ModSwitchHandler (ModSwitch.cu:175-232 with level -1: beta 1,
// no padding, GetSSENumAux() = num_q auxiliary primes) and P' mod q for its
// fold (Context.cpp:109-123). The short base must be the low end of the
// bottom level (its terminal primes the last ones, its main primes the
// first ones), with at most alpha primes.
int build_short_base(Context &c, int num_main, int num_ter) {
  const int num_q = num_main + num_ter;
  if (num_main < 0 || num_ter < 0 || num_q < 1 || num_q > c.alpha ||
      num_main > c.lv_main[0] || num_ter > c.lv_ter[0]) {
    set_error("short base must be 1..alpha primes within level 0");
    return kErrInvalidArg;
  }
  if (c.short_naux > 0) {
    if (c.short_base.nm == num_main && c.short_base.nt == num_ter) return kOk;
    set_error("a different short base is already set up");
    return kErrInvalidArg;
  }
  LevelPlan lp;
  int rc = build_modup(c, lp, num_main, num_ter, num_q, 0);
  if (!rc) rc = build_moddown(c, num_main, num_ter, num_main, num_ter, num_q, 0, lp.moddown);
  if (rc) return rc;
  std::vector<uint32_t> pp(c.num_primes, 0);
  for (int t = 0; t < num_q; t++) {
    const int p = lp.q_prime[t];
    const uint32_t q = c.primes[p];
    uint32_t v = 1;
    for (int j = 0; j < num_q; j++) v = h_mulmod(v, c.primes[c.L + j] % q, q);
    pp[p] = h_to_mont(v, q);
  }
  rc = upload(c, &c.d_p_prod_short, pp);
  if (rc) return rc;
  c.short_base = std::move(lp);
  c.short_naux = num_q;
  return kOk;
}

int context_create(int log_n, const uint32_t *main_primes, int n_main,
                   const uint32_t *ter_primes, int n_ter,
                   const uint32_t *aux_primes, int n_aux,
                   const int *level_cfg, int n_levels, Context **out) {
  if (!out) {
    set_error("out is null");
    return kErrInvalidArg;
  }
  *out = nullptr;
  if (log_n != 16) {
    set_error("only log_degree 16 is supported on gfx950");
    return kErrUnsupported;
  }
  if (n_main < 0 || n_ter < 0 || n_aux <= 0 || n_main + n_ter <= 0 ||
      n_levels <= 0 || !level_cfg || (n_main && !main_primes) ||
      (n_ter && !ter_primes) || !aux_primes) {
    set_error("invalid parameter arrays");
    return kErrInvalidArg;
  }
  const int n = 1 << log_n;
  const int np = n_main + n_ter + n_aux;
  if (np > 255) {
    set_error("at most 255 primes are supported");
    return kErrUnsupported;
  }
  std::vector<uint32_t> all;
  for (int i = n_ter - 1; i >= 0; i--) all.push_back(ter_primes[i]);
  for (int i = 0; i < n_main; i++) all.push_back(main_primes[i]);
  for (int i = 0; i < n_aux; i++) all.push_back(aux_primes[i]);
  for (uint32_t q : all) {
    // Parameter.cu:43-64
    if (q >= (1u << 31) || q % (2u * n) != 1 || !is_prime_u32(q)) {
      set_error("prime " + std::to_string(q) +
                " is not an NTT-friendly prime below 2^31");
      return kErrInvalidArg;
    }
  }
  // Parameter.cu:66-91
  int pm = 0, pt = 0;
  for (int i = 0; i < n_levels - 1; i++) {
    int m = level_cfg[2 * i], t = level_cfg[2 * i + 1];
    if (m < 0 || m > n_main || t < 0 || t > n_ter || !np_less(pm, pt, m, t)) {
      set_error("invalid level configuration at level " + std::to_string(i));
      return kErrInvalidArg;
    }
    pm = m;
    pt = t;
  }
  if (level_cfg[2 * (n_levels - 1)] != n_main ||
      level_cfg[2 * (n_levels - 1) + 1] != n_ter) {
    set_error("the last level must hold every main and terminal prime");
    return kErrInvalidArg;
  }

  Context *c = new Context();
  c->log_n = log_n;
  c->n = n;
  c->max_main = n_main;
  c->max_ter = n_ter;
  c->alpha = n_aux;
  c->L = n_main + n_ter;
  c->dnum = (c->L + n_aux - 1) / n_aux;
  c->max_level = n_levels - 1;
  c->num_primes = np;
  c->primes = all;
  for (int i = 0; i < n_levels; i++) {
    c->lv_main.push_back(level_cfg[2 * i]);
    c->lv_ter.push_back(level_cfg[2 * i + 1]);
  }
  if (c->dnum > 16) {
    delete c;
    set_error("dnum above 16 is not supported");
    return kErrUnsupported;
  }

  std::vector<int32_t> qinv(np);
  std::vector<double> qd(np), qr(np);
  std::vector<uint32_t> tw_f((size_t)np * n), tw_i((size_t)np * n), imag(np);
  c->psi.assign(np, 0);
  for (int i = 0; i < np; i++) {
    const uint32_t q = all[i];
    qinv[i] = inv_mod_2_32(q);
    qd[i] = (double)q;
    qr[i] = 1.0 / (double)q;
    const uint32_t g = smallest_generator(q);
    const uint32_t psi = h_powmod(g, (q - 1) / (2u * n), q);
    const uint32_t psi_inv = h_invmod(psi, q);
    c->psi[i] = psi;
    uint64_t pf = 1, pi = 1;
    for (int j = 0; j < n; j++) {
      const uint32_t r = brv((uint32_t)j, log_n);
      tw_f[(size_t)i * n + r] = (uint32_t)pf;
      tw_i[(size_t)i * n + r] = (uint32_t)pi;
      pf = pf * psi % q;
      pi = pi * psi_inv % q;
    }
    imag[i] = h_to_mont(h_powmod(psi, n / 2, q), q);
  }
  c->qinv = qinv;
  // P mod q_i in Montgomery form (Context.cpp:85-98)
  c->p_prod_mont.assign(np, 0);
  for (int i = 0; i < c->L; i++) {
    uint32_t v = 1;
    for (int a = 0; a < n_aux; a++)
      v = h_mulmod(v, all[c->L + a] % all[i], all[i]);
    c->p_prod_mont[i] = h_to_mont(v, all[i]);
  }
  // Galois factors 5^i mod 2N (Parameter.cu:93-101)
  const int slots = n / 2;
  c->galois.assign(slots + 1, 1);
  for (int i = 1; i < slots; i++)
    c->galois[i] = (uint32_t)((uint64_t)c->galois[i - 1] * 5u % (2u * n));

  int rc = upload(*c, &c->d_q, all);
  if (!rc) rc = upload(*c, &c->d_qinv, qinv);
  if (!rc) rc = upload(*c, &c->d_qd, qd);
  if (!rc) rc = upload(*c, &c->d_qr, qr);
  if (!rc) rc = upload(*c, &c->d_tw_fwd, tw_f);
  if (!rc) rc = upload(*c, &c->d_tw_inv, tw_i);
  if (!rc) rc = upload(*c, &c->d_p_prod, c->p_prod_mont);
  if (!rc) rc = upload(*c, &c->d_imag, imag);
  if (rc) {
    context_destroy(c);
    return rc;
  }
  c->ptab = PrimeTab{c->d_q, c->d_qinv, c->d_qd, c->d_qr};

  c->levels.resize(n_levels);
  for (int l = 0; l < n_levels; l++) {
    rc = build_level(*c, l);
    if (rc) {
      context_destroy(c);
      return rc;
    }
  }
  // workspace for the largest operation at any level: HRot (its MultKey
  // output + a relinearization's temporaries) or up to kMaxBsRot hoisted
  // rotations with their ModDown
  size_t limbs = 0;
  for (int l = 0; l < n_levels; l++) {
    const size_t hrot = 2 * (size_t)c->levels[l].num_q + relin_ws_limbs(*c, l);
    limbs = std::max({limbs, hrot, hoisted_ws_limbs(*c, l, kMaxBsRot, true)});
  }
  c->ws_words = limbs * (size_t)n;
  rc = check_hip(hipMalloc((void **)&c->ws, c->ws_words * sizeof(uint32_t)),
                 "hipMalloc(workspace)");
  if (rc) {
    context_destroy(c);
    return rc;
  }
  c->allocations.push_back(c->ws);
  *out = c;
  return kOk;
}

void prof_free(const Context &c);

void context_destroy(Context *ctx) {
  if (!ctx) return;
  prof_free(*ctx);
  for (auto &e : ctx->i8_bconv_planes) (void)hipFree(e.second);
  for (void *p : ctx->allocations) (void)hipFree(p);
  delete ctx;
}

}  // namespace ck

namespace ck {

namespace {
hipEvent_t pooled_event(const Context &c) {
  if (!c.prof_pool.empty()) {
    hipEvent_t e = c.prof_pool.back();
    c.prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

bool prof_events(const Context &c, const char *kernel, int limbs,
                 hipEvent_t *start, hipEvent_t *stop) {
  if (!c.prof_on) return false;
  const std::string key = std::string(kernel) + "/" + std::to_string(limbs);
  int id = -1;
  for (size_t i = 0; i < c.prof_names.size(); i++)
    if (c.prof_names[i] == key) id = (int)i;
  if (id < 0) {
    id = (int)c.prof_names.size();
    c.prof_names.push_back(key);
  }
  Context::ProfEvent ev{id, pooled_event(c), pooled_event(c)};
  if (!ev.start || !ev.stop) return false;
  c.prof_events.push_back(ev);
  *start = ev.start;
  *stop = ev.stop;
  return true;
}

// Accumulate the recorded launches per kernel name (waits for the events).
int prof_report(const Context &c, std::vector<std::string> &names,
                std::vector<double> &total_ms, std::vector<long> &calls) {
  names = c.prof_names;
  total_ms.assign(names.size(), 0.0);
  calls.assign(names.size(), 0);
  for (const auto &ev : c.prof_events) {
    int rc = check_hip(hipEventSynchronize(ev.stop), "profile sync");
    if (rc) return rc;
    float ms = 0.f;
    rc = check_hip(hipEventElapsedTime(&ms, ev.start, ev.stop), "profile time");
    if (rc) return rc;
    total_ms[ev.kernel] += ms;
    calls[ev.kernel] += 1;
  }
  return kOk;
}

void prof_reset(const Context &c) {
  for (const auto &ev : c.prof_events) {
    c.prof_pool.push_back(ev.start);
    c.prof_pool.push_back(ev.stop);
  }
  c.prof_events.clear();
  c.prof_names.clear();
}

void prof_free(const Context &c) {
  prof_reset(c);
  for (hipEvent_t e : c.prof_pool) (void)hipEventDestroy(e);
  c.prof_pool.clear();
}

}  // namespace ck
