// extern "C" entry points of include/cheddar_amd.h: argument validation
// (the reference's AssertTrue checks, returned as status codes) and
// dispatch to the host orchestration in ck_ops.cpp.
#include "../../include/cheddar_amd.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "ck_plan.h"

// Workspace of ck_hmult_noalloc: a file-scope device array (192 limbs of
// N = 2^16 words, 48 MiB). Every device the library runs on holds its own
// copy from module load until the process ends. Host code never uses the
// array's name as an address: scratch_ptr() returns the device address that
// hipGetSymbolAddress reports for the current device.
namespace {
constexpr size_t kScratchLimbs = 192;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t *g_scratch_ptr[kMaxDevices] = {};
}  // namespace
static __device__ __align__(256) uint32_t scratch[kScratchLimbs * 65536];

struct ck_context {
  ck::Context *c;
};

struct ck_plan {
  ck::Plan *p;
};

using namespace ck;

namespace {

constexpr size_t kLimb = 65536;

int fail(int code, const char *msg) {
  set_error(msg);
  return code;
}

#define CK_REQUIRE(cond, msg) \
  do {                        \
    if (!(cond)) return fail(CK_ERR_INVALID_ARG, msg); \
  } while (0)

hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept
int scratch_ptr(uint32_t **out) {
  *out = nullptr;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(CK_ERR_HIP, "no current HIP device for the static scratch");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return CK_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<uint32_t *>(p);
  }
  *out = g_scratch_ptr[dev];
  return CK_OK;
}

bool valid_np(const Context &c, int nm, int nt, int na) {
  return nm >= 0 && nm <= c.max_main && nt >= 0 && nt <= c.max_ter &&
         na >= 0 && na <= c.alpha && nm + nt + na > 0;
}

bool valid_level(const Context &c, int level) {
  return level >= 0 && level <= c.max_level;
}

bool valid_ptrs(const uint32_t *const *p, int n) {
  if (!p) return false;
  for (int i = 0; i < n; i++)
    if (!p[i]) return false;
  return true;
}

bool valid_key(const Context &c, const uint32_t *const *k, int level) {
  if (!k) return false;
  const LevelPlan &lp = level_plan(c, level);
  for (int g = 0; g < lp.beta; g++)
    if (lp.groups[g].s1 > lp.groups[g].s0 && !k[g]) return false;
  return true;
}

int elementwise(ck_context *ctx, int op, int nm, int nt, int na, int npoly,
                uint32_t *const *dst, const uint32_t *const *x,
                const uint32_t *const *y, void *stream) {
  CK_REQUIRE(ctx && ctx->c, "null context");
  const Context &c = *ctx->c;
  CK_REQUIRE(valid_np(c, nm, nt, na), "invalid NP");
  CK_REQUIRE(npoly >= 1 && npoly <= kMaxPolys, "npoly must be 1..3");
  CK_REQUIRE(valid_ptrs(dst, npoly) && valid_ptrs(x, npoly), "null pointer");
  EwNp np = make_np(c, nm, nt, na);
  EwOperand ox{}, oy{}, od{};
  for (int j = 0; j < npoly; j++) {
    ox.ptr[j] = const_cast<uint32_t *>(x[j]);
    od.ptr[j] = dst[j];
  }
  if (op <= 2) {
    CK_REQUIRE(valid_ptrs(y, npoly), "null pointer");
    for (int j = 0; j < npoly; j++) oy.ptr[j] = const_cast<uint32_t *>(y[j]);
    launch_binary(c, op, np, ox, oy, od, npoly, as_stream(stream));
  } else {
    launch_unary(c, 0, np, ox, od, npoly, as_stream(stream));
  }
  return check_hip(hipGetLastError(), "elementwise");
}

}  // namespace

extern "C" {

const char *ck_last_error(void) { return ck::last_error(); }

int ck_version(void) { return 100; }

int ck_context_create(int log_degree, const uint32_t *main_primes, int n_main,
                      const uint32_t *ter_primes, int n_ter,
                      const uint32_t *aux_primes, int n_aux,
                      const int *level_config, int n_levels,
                      ck_context **out) {
  CK_REQUIRE(out, "out is null");
  *out = nullptr;
  Context *c = nullptr;
  int rc = context_create(log_degree, main_primes, n_main, ter_primes, n_ter,
                          aux_primes, n_aux, level_config, n_levels, &c);
  if (rc) return rc;
  // set-up of ck_hmult_noalloc's static scratch on this device
  uint32_t *sp = nullptr;
  rc = scratch_ptr(&sp);
  if (rc) {
    context_destroy(c);
    return rc;
  }
  *out = new ck_context{c};
  return CK_OK;
}

int ck_context_destroy(ck_context *ctx) {
  if (!ctx) return CK_OK;
  context_destroy(ctx->c);
  delete ctx;
  return CK_OK;
}

int ck_context_info(const ck_context *ctx, int *out) {
  CK_REQUIRE(ctx && ctx->c && out, "null argument");
  const Context &c = *ctx->c;
  out[0] = c.log_n;
  out[1] = c.n;
  out[2] = c.max_main;
  out[3] = c.max_ter;
  out[4] = c.alpha;
  out[5] = c.L;
  out[6] = c.dnum;
  out[7] = c.max_level;
  out[8] = c.num_primes;
  out[9] = (int)((c.ws_words * sizeof(uint32_t)) >> 20);
  return CK_OK;
}

int ck_level_to_np(const ck_context *ctx, int level, int *num_main,
                   int *num_ter) {
  CK_REQUIRE(ctx && ctx->c && num_main && num_ter, "null argument");
  CK_REQUIRE(has_switch_level(*ctx->c, level), "invalid level");
  const LevelPlan &lp = level_plan(*ctx->c, level);
  *num_main = lp.nm;
  *num_ter = lp.nt;
  return CK_OK;
}

int ck_np_to_level(const ck_context *ctx, int num_main, int num_ter) {
  if (!ctx || !ctx->c) return -1;
  const Context &c = *ctx->c;
  for (int l = 0; l <= c.max_level; l++)
    if (c.lv_main[l] == num_main && c.lv_ter[l] == num_ter) return l;
  return -1;
}

int ck_context_set_short_base(ck_context *ctx, int num_main, int num_ter) {
  CK_REQUIRE(ctx && ctx->c, "null context");
  return build_short_base(*ctx->c, num_main, num_ter);
}

int ck_modup_beta(const ck_context *ctx, int level) {
  if (!ctx || !ctx->c || !has_switch_level(*ctx->c, level)) return -1;
  return level_plan(*ctx->c, level).beta;
}

int ck_galois_factor(const ck_context *ctx, int rot, uint32_t *out) {
  CK_REQUIRE(ctx && ctx->c && out, "null argument");
  const Context &c = *ctx->c;
  if (rot == -1) {
    *out = 2u * c.n - 1;
    return CK_OK;
  }
  CK_REQUIRE(rot >= 0 && rot <= c.n / 2, "invalid permute amount");
  *out = c.galois[rot];
  return CK_OK;
}

int ck_prime(const ck_context *ctx, int pidx, uint32_t *out) {
  CK_REQUIRE(ctx && ctx->c && out, "null argument");
  CK_REQUIRE(pidx >= 0 && pidx < ctx->c->num_primes, "invalid prime index");
  *out = ctx->c->primes[pidx];
  return CK_OK;
}

int ck_ntt(ck_context *ctx, uint32_t *dst, const uint32_t *src, int num_main,
           int num_ter, int num_aux, int src_gap, int montgomery_conversion,
           void *stream) {
  CK_REQUIRE(ctx && ctx->c && dst && src, "null argument");
  CK_REQUIRE(valid_np(*ctx->c, num_main, num_ter, num_aux), "invalid NP");
  CK_REQUIRE(src_gap >= 0 && (src_gap == 0 || src != dst),
             "a limb gap requires dst != src");
  ImmediateSink sink(*ctx->c, as_stream(stream));
  return op_ntt(*ctx->c, dst, src, num_main, num_ter, num_aux, src_gap,
                montgomery_conversion != 0, sink);
}

int ck_intt(ck_context *ctx, uint32_t *dst, const uint32_t *src, int num_main,
            int num_ter, int num_aux, int src_gap, int montgomery_conversion,
            void *stream) {
  CK_REQUIRE(ctx && ctx->c && dst && src, "null argument");
  CK_REQUIRE(valid_np(*ctx->c, num_main, num_ter, num_aux), "invalid NP");
  CK_REQUIRE(src_gap >= 0 && (src_gap == 0 || src != dst),
             "a limb gap requires dst != src");
  ImmediateSink sink(*ctx->c, as_stream(stream));
  return op_intt(*ctx->c, dst, src, num_main, num_ter, num_aux, src_gap,
                 montgomery_conversion != 0, sink);
}

int ck_intt_mult_const(ck_context *ctx, uint32_t *dst, const uint32_t *src,
                       int num_main, int num_ter, int num_aux, int src_gap,
                       const uint32_t *consts, int normalize, void *stream) {
  CK_REQUIRE(ctx && ctx->c && dst && src && consts, "null argument");
  CK_REQUIRE(valid_np(*ctx->c, num_main, num_ter, num_aux), "invalid NP");
  CK_REQUIRE(src_gap >= 0 && (src_gap == 0 || src != dst),
             "a limb gap requires dst != src");
  ImmediateSink sink(*ctx->c, as_stream(stream));
  return op_intt_const(*ctx->c, dst, src, num_main, num_ter, num_aux, src_gap,
                       consts, normalize != 0, sink);
}

int ck_modup(ck_context *ctx, int level, uint32_t *dst, const uint32_t *src,
             void *stream) {
  CK_REQUIRE(ctx && ctx->c && dst && src, "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(has_switch_level(c, level), "invalid level");
  CK_REQUIRE(dst != src, "ModUp is not in place");
  ImmediateSink sink(c, as_stream(stream));
  return op_modup(c, level, dst, src, c.ws, sink);
}

int ck_moddown(ck_context *ctx, int level, int kind, int npoly,
               uint32_t *const *dst, const uint32_t *const *src,
               void *stream) {
  CK_REQUIRE(ctx && ctx->c, "null context");
  const Context &c = *ctx->c;
  CK_REQUIRE(has_switch_level(c, level), "invalid level");
  CK_REQUIRE(kind >= 0 && kind <= 2, "invalid ModDown kind");
  CK_REQUIRE(npoly >= 1 && npoly <= 4, "npoly must be 1..4");
  CK_REQUIRE(valid_ptrs(dst, npoly) && valid_ptrs(src, npoly), "null pointer");
  for (int j = 0; j < npoly; j++)
    for (int k = 0; k < npoly; k++)
      CK_REQUIRE(dst[j] != src[k], "ModDown dst and src must differ");
  CK_REQUIRE(moddown_plan(c, level, kind), "no such rescale at this level");
  ImmediateSink sink(c, as_stream(stream));
  return op_moddown(c, level, kind, npoly, dst, src, c.ws, sink);
}

int ck_copy(ck_context *ctx, void *dst, const void *src, size_t bytes, void *stream) {
  CK_REQUIRE(ctx && ctx->c, "null context");
  if (bytes == 0) return CK_OK;
  CK_REQUIRE(dst && src, "null pointer");
  const char *d = static_cast<const char *>(dst), *s = static_cast<const char *>(src);
  CK_REQUIRE(d == s || d + bytes <= s || s + bytes <= d, "copy ranges overlap");
  return launch_copy(*ctx->c, dst, src, bytes, as_stream(stream));
}

int ck_add(ck_context *ctx, int num_main, int num_ter, int num_aux, int npoly,
           uint32_t *const *dst, const uint32_t *const *x,
           const uint32_t *const *y, void *stream) {
  return elementwise(ctx, 0, num_main, num_ter, num_aux, npoly, dst, x, y,
                     stream);
}

int ck_sub(ck_context *ctx, int num_main, int num_ter, int num_aux, int npoly,
           uint32_t *const *dst, const uint32_t *const *x,
           const uint32_t *const *y, void *stream) {
  return elementwise(ctx, 1, num_main, num_ter, num_aux, npoly, dst, x, y,
                     stream);
}

int ck_mult(ck_context *ctx, int num_main, int num_ter, int num_aux,
            int npoly, uint32_t *const *dst, const uint32_t *const *x,
            const uint32_t *const *y, void *stream) {
  return elementwise(ctx, 2, num_main, num_ter, num_aux, npoly, dst, x, y,
                     stream);
}

int ck_neg(ck_context *ctx, int num_main, int num_ter, int num_aux, int npoly,
           uint32_t *const *dst, const uint32_t *const *x, void *stream) {
  return elementwise(ctx, 3, num_main, num_ter, num_aux, npoly, dst, x,
                     nullptr, stream);
}

int ck_mult_const(ck_context *ctx, int num_main, int num_ter, int num_aux,
                  int npoly, uint32_t *const *dst, const uint32_t *const *x,
                  const uint32_t *consts, void *stream) {
  CK_REQUIRE(ctx && ctx->c && consts, "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(valid_np(c, num_main, num_ter, num_aux), "invalid NP");
  CK_REQUIRE(npoly >= 1 && npoly <= kMaxPolys, "npoly must be 1..3");
  CK_REQUIRE(valid_ptrs(dst, npoly) && valid_ptrs(x, npoly), "null pointer");
  EwNp np = make_np(c, num_main, num_ter, num_aux);
  EwOperand ox{}, od{};
  for (int j = 0; j < npoly; j++) {
    ox.ptr[j] = const_cast<uint32_t *>(x[j]);
    od.ptr[j] = dst[j];
  }
  launch_const(c, 0, np, ox, consts, 0, od, npoly, as_stream(stream));
  return check_hip(hipGetLastError(), "mult_const");
}

int ck_add_const(ck_context *ctx, int num_main, int num_ter, int num_aux,
                 int npoly, uint32_t *const *dst, const uint32_t *const *x,
                 const uint32_t *consts, int mode, void *stream) {
  CK_REQUIRE(ctx && ctx->c && consts, "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(valid_np(c, num_main, num_ter, num_aux), "invalid NP");
  CK_REQUIRE(npoly >= 1 && npoly <= kMaxPolys, "npoly must be 1..3");
  CK_REQUIRE(mode >= 0 && mode <= 2, "mode must be 0 (add), 1 (sub) or 2 (c - x)");
  CK_REQUIRE(valid_ptrs(dst, npoly) && valid_ptrs(x, npoly), "null pointer");
  EwNp np = make_np(c, num_main, num_ter, num_aux);
  EwOperand ox{}, od{};
  for (int j = 0; j < npoly; j++) {
    ox.ptr[j] = const_cast<uint32_t *>(x[j]);
    od.ptr[j] = dst[j];
  }
  launch_const(c, mode + 1, np, ox, consts, 0, od, npoly, as_stream(stream));
  return check_hip(hipGetLastError(), "add_const");
}

int ck_mult_imaginary_unit(ck_context *ctx, int num_main, int num_ter,
                           int num_aux, int npoly, uint32_t *const *dst,
                           const uint32_t *const *x, void *stream) {
  CK_REQUIRE(ctx && ctx->c, "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(valid_np(c, num_main, num_ter, num_aux), "invalid NP");
  CK_REQUIRE(npoly >= 1 && npoly <= kMaxPolys, "npoly must be 1..3");
  CK_REQUIRE(valid_ptrs(dst, npoly) && valid_ptrs(x, npoly), "null pointer");
  EwNp np = make_np(c, num_main, num_ter, num_aux);
  EwOperand ox{}, od{};
  for (int j = 0; j < npoly; j++) {
    ox.ptr[j] = const_cast<uint32_t *>(x[j]);
    od.ptr[j] = dst[j];
  }
  launch_imag_unit(c, np, ox, od, npoly, as_stream(stream));
  return check_hip(hipGetLastError(), "mult_imaginary_unit");
}

int ck_modup_to_max(ck_context *ctx, int base_main, int base_ter,
                    const uint32_t *src, uint32_t *dst, int num_aux,
                    void *stream) {
  CK_REQUIRE(ctx && ctx->c && src && dst, "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(valid_np(c, base_main, base_ter, 0), "invalid base NP");
  const int nbase = base_main + base_ter;
  CK_REQUIRE(nbase == 1 || nbase == 2, "the base must hold one or two primes");
  CK_REQUIRE(num_aux >= 0 && num_aux <= c.alpha, "invalid num_aux");
  CK_REQUIRE(dst != src, "dst must not alias src");
  const int p0 = c.max_ter - base_ter;
  LiftArgs A{};
  A.nsrc = nbase;
  A.q0 = c.primes[p0];
  A.q1 = nbase == 2 ? c.primes[p0 + 1] : 0u;
  const int ndst = c.L + num_aux;  // every q prime (global order), then aux
  CK_REQUIRE(ndst <= kMaxLiftDst, "too many destination limbs");
  for (int y = 0; y < ndst; y++) A.dst_prime[y] = (uint16_t)y;
  launch_crt_lift(c, src, A, ndst, dst, as_stream(stream));
  return check_hip(hipGetLastError(), "modup_to_max");
}

int ck_crt_lift(ck_context *ctx, const int *src_primes, int n_src,
                const int *dst_primes, int n_dst, const uint32_t *src, uint32_t *dst,
                void *stream) {
  CK_REQUIRE(ctx && ctx->c && src_primes && dst_primes && src && dst, "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(n_src == 1 || n_src == 2, "one or two source primes");
  CK_REQUIRE(n_dst >= 1 && n_dst <= kMaxLiftDst, "1..128 destination primes");
  CK_REQUIRE(dst != src, "dst must not alias src");
  const int np = (int)c.primes.size();
  for (int i = 0; i < n_src; i++)
    CK_REQUIRE(src_primes[i] >= 0 && src_primes[i] < np, "invalid source prime");
  CK_REQUIRE(n_src == 1 || src_primes[0] != src_primes[1], "source primes must differ");
  LiftArgs A{};
  A.nsrc = n_src;
  A.q0 = c.primes[src_primes[0]];
  A.q1 = n_src == 2 ? c.primes[src_primes[1]] : 0u;
  for (int y = 0; y < n_dst; y++) {
    CK_REQUIRE(dst_primes[y] >= 0 && dst_primes[y] < np, "invalid destination prime");
    A.dst_prime[y] = (uint16_t)dst_primes[y];
  }
  launch_crt_lift(c, src, A, n_dst, dst, as_stream(stream));
  return check_hip(hipGetLastError(), "crt_lift");
}

static int accum_common(ck_context *ctx, int num_main, int num_ter, int num_aux,
                        int npoly, int nterms, uint32_t *const *dst,
                        const uint32_t *const *x, const uint32_t *const *y,
                        int y_const, int accumulate, void *stream) {
  CK_REQUIRE(ctx && ctx->c && dst && (nterms == 0 || (x && y)), "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(valid_np(c, num_main, num_ter, num_aux), "invalid NP");
  CK_REQUIRE(npoly >= 1 && npoly <= 2, "npoly must be 1 or 2");
  CK_REQUIRE(nterms >= 0, "nterms must not be negative");
  CK_REQUIRE(valid_ptrs(dst, npoly), "null pointer");
  CK_REQUIRE(nterms == 0 || (valid_ptrs(x, nterms * npoly) && valid_ptrs(y, nterms)),
             "null pointer");
  EwNp np = make_np(c, num_main, num_ter, num_aux);
  int done = 0;
  do {
    AccumArgs A{};
    A.nterms = std::min(nterms - done, kMaxAccTerms);
    A.npoly = npoly;
    A.y_const = y_const;
    A.accumulate = accumulate || done > 0;
    for (int j = 0; j < npoly; j++) A.dst[j] = dst[j];
    for (int t = 0; t < A.nterms; t++) {
      for (int j = 0; j < npoly; j++) A.x[t][j] = x[(size_t)(done + t) * npoly + j];
      A.y[t] = y[done + t];
    }
    launch_accum(c, np, A, as_stream(stream));
    done += A.nterms;
  } while (done < nterms);
  return check_hip(hipGetLastError(), y_const ? "caccum" : "paccum");
}

int ck_paccum(ck_context *ctx, int num_main, int num_ter, int num_aux, int npoly,
              int nterms, uint32_t *const *dst, const uint32_t *const *x,
              const uint32_t *const *y, int accumulate, void *stream) {
  return accum_common(ctx, num_main, num_ter, num_aux, npoly, nterms, dst, x, y, 0,
                      accumulate, stream);
}

int ck_caccum(ck_context *ctx, int num_main, int num_ter, int num_aux, int npoly,
              int nterms, uint32_t *const *dst, const uint32_t *const *x,
              const uint32_t *const *consts, int accumulate, void *stream) {
  return accum_common(ctx, num_main, num_ter, num_aux, npoly, nterms, dst, x, consts, 1,
                      accumulate, stream);
}

int ck_gs_accum(ck_context *ctx, int num_main, int num_ter, int num_aux, int nbs,
                int ngs, uint32_t *const *dst, const uint32_t *const *x,
                const uint32_t *const *pt, void *stream) {
  CK_REQUIRE(ctx && ctx->c && dst && x && pt, "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(valid_np(c, num_main, num_ter, num_aux), "invalid NP");
  CK_REQUIRE(nbs >= 1 && nbs <= kMaxGsBaby, "nbs must be 1..16");
  CK_REQUIRE(ngs >= 0, "ngs must not be negative");
  CK_REQUIRE(valid_ptrs(x, 2 * nbs) && valid_ptrs(dst, 2 * ngs), "null pointer");
  // an output aliasing an operand would be read after it was written
  for (int d = 0; d < 2 * ngs; d++) {
    for (int t = 0; t < 2 * nbs; t++)
      CK_REQUIRE(dst[d] != x[t], "gs_accum: an output aliases a baby step");
    for (int t = 0; t < ngs * nbs; t++)
      CK_REQUIRE(dst[d] != pt[t], "gs_accum: an output aliases a plaintext");
  }
  const EwNp np = make_np(c, num_main, num_ter, num_aux);
  for (int g0 = 0; g0 < ngs; g0 += kMaxGsGiant) {
    GsAccumArgs A{};
    A.nbs = nbs;
    A.ngs = std::min(ngs - g0, kMaxGsGiant);
    for (int t = 0; t < nbs; t++) {
      A.x[t][0] = x[2 * t];
      A.x[t][1] = x[2 * t + 1];
    }
    for (int g = 0; g < A.ngs; g++) {
      A.dst[g][0] = dst[2 * (g0 + g)];
      A.dst[g][1] = dst[2 * (g0 + g) + 1];
      for (int t = 0; t < nbs; t++) A.pt[g][t] = pt[(size_t)(g0 + g) * nbs + t];
    }
    launch_gs_accum(c, np, A, as_stream(stream));
  }
  return check_hip(hipGetLastError(), "gs_accum");
}

int ck_permute_accum(ck_context *ctx, int num_main, int num_ter, int num_aux, int npoly,
                     uint32_t *const *dst, int nsrc, const uint32_t *const *src,
                     const int *rots, int accumulate, void *stream) {
  CK_REQUIRE(ctx && ctx->c && dst && (nsrc == 0 || (src && rots)), "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(valid_np(c, num_main, num_ter, num_aux), "invalid NP");
  CK_REQUIRE(npoly >= 1 && npoly <= 2, "npoly must be 1 or 2");
  CK_REQUIRE(nsrc >= 0, "nsrc must not be negative");
  CK_REQUIRE(valid_ptrs(dst, npoly), "null pointer");
  for (int s = 0; s < nsrc * npoly; s++)
    for (int j = 0; j < npoly; j++)
      CK_REQUIRE(src[s] == nullptr || src[s] != dst[j],
                 "PermuteAccum does not support a source aliasing the output");
  // no sources: accumulating leaves dst as it is; otherwise one pass of
  // masked terms writes zeros
  if (nsrc == 0 && accumulate) return CK_OK;
  const EwNp np = make_np(c, num_main, num_ter, num_aux);
  int done = 0;
  do {
    PermAccArgs A{};
    A.npoly = npoly;
    A.nsrc = std::min(nsrc - done, kMaxPermAcc);
    A.accumulate = accumulate || done > 0;
    for (int j = 0; j < npoly; j++) A.dst[j] = dst[j];
    for (int s = 0; s < A.nsrc; s++) {
      uint32_t g = 1;
      if (rots[done + s] != 0) {
        int rc = ck_galois_factor(ctx, rots[done + s], &g);
        if (rc) return rc;
      }
      A.g[s] = g;
      for (int j = 0; j < npoly; j++) A.src[s][j] = src[(size_t)(done + s) * npoly + j];
    }
    launch_permute_accum(c, np, A, as_stream(stream));
    done += A.nsrc;
  } while (done < nsrc);
  return check_hip(hipGetLastError(), "permute_accum");
}

int ck_tensor(ck_context *ctx, int num_main, int num_ter, uint32_t *d0,
              uint32_t *d1, uint32_t *d2, const uint32_t *b1,
              const uint32_t *a1, const uint32_t *b2, const uint32_t *a2,
              void *stream) {
  CK_REQUIRE(ctx && ctx->c && d0 && d1 && d2 && b1 && a1 && b2 && a2,
             "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(valid_np(c, num_main, num_ter, 0), "invalid NP");
  EwNp np = make_np(c, num_main, num_ter, 0);
  EwOperand x{{const_cast<uint32_t *>(b1), const_cast<uint32_t *>(a1), nullptr},
              0};
  EwOperand y{{const_cast<uint32_t *>(b2), const_cast<uint32_t *>(a2), nullptr},
              0};
  EwOperand o{{d0, d1, d2}, 0};
  launch_tensor(c, np, x, y, o, as_stream(stream));
  return check_hip(hipGetLastError(), "tensor");
}

int ck_permute(ck_context *ctx, int num_main, int num_ter, int num_aux,
               int npoly, uint32_t *const *dst, const uint32_t *const *x,
               int rot, void *stream) {
  CK_REQUIRE(ctx && ctx->c, "null context");
  const Context &c = *ctx->c;
  CK_REQUIRE(valid_np(c, num_main, num_ter, num_aux), "invalid NP");
  CK_REQUIRE(npoly >= 1 && npoly <= kMaxPolys, "npoly must be 1..3");
  CK_REQUIRE(valid_ptrs(dst, npoly) && valid_ptrs(x, npoly), "null pointer");
  for (int j = 0; j < npoly; j++)
    CK_REQUIRE(dst[j] != x[j], "Permute does not support in-place operation");
  uint32_t g = 0;
  int rc = ck_galois_factor(ctx, rot, &g);
  if (rc) return rc;
  EwNp np = make_np(c, num_main, num_ter, num_aux);
  EwOperand ox{}, od{};
  for (int j = 0; j < npoly; j++) {
    ox.ptr[j] = const_cast<uint32_t *>(x[j]);
    od.ptr[j] = dst[j];
  }
  launch_permute(c, np, ox, od, g, npoly, as_stream(stream));
  return check_hip(hipGetLastError(), "permute");
}

int ck_mult_key_no_moddown(ck_context *ctx, int level, uint32_t *accum_b,
                           uint32_t *accum_a, const uint32_t *x,
                           const uint32_t *add_b, const uint32_t *add_a,
                           const uint32_t *const *key_b,
                           const uint32_t *const *key_a, void *stream) {
  CK_REQUIRE(ctx && ctx->c && accum_b && accum_a && x, "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(has_switch_level(c, level), "invalid level");
  CK_REQUIRE(valid_key(c, key_b, level) && valid_key(c, key_a, level),
             "null key block");
  CK_REQUIRE(accum_b != x && accum_a != x,
             "in-place operation is not supported for MultKeyNoModDown");
  ImmediateSink sink(c, as_stream(stream));
  return op_mult_key_no_moddown(c, level, accum_b, accum_a, x, add_b, add_a,
                                key_b, key_a, c.ws, sink);
}

int ck_mult_key(ck_context *ctx, int level, uint32_t *out_b, uint32_t *out_a,
                const uint32_t *b, const uint32_t *a,
                const uint32_t *const *key_b, const uint32_t *const *key_a,
                void *stream) {
  CK_REQUIRE(ctx && ctx->c && out_b && out_a && b && a, "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(has_switch_level(c, level), "invalid level");
  CK_REQUIRE(valid_key(c, key_b, level) && valid_key(c, key_a, level),
             "null key block");
  ImmediateSink sink(c, as_stream(stream));
  return op_mult_key(c, level, out_b, out_a, b, a, key_b, key_a, c.ws, sink);
}

int ck_relinearize(ck_context *ctx, int level, uint32_t *out_b,
                   uint32_t *out_a, const uint32_t *d0, const uint32_t *d1,
                   const uint32_t *d2, const uint32_t *const *key_b,
                   const uint32_t *const *key_a, int rescale, void *stream) {
  CK_REQUIRE(ctx && ctx->c && out_b && out_a && d0 && d1 && d2,
             "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(valid_level(c, level), "invalid level");
  CK_REQUIRE(!rescale || moddown_plan(c, level, 2),
             "cannot rescale at this level");
  CK_REQUIRE(valid_key(c, key_b, level) && valid_key(c, key_a, level),
             "null key block");
  ImmediateSink sink(c, as_stream(stream));
  return op_relinearize(c, level, out_b, out_a, d0, d1, d2, key_b, key_a,
                        rescale != 0, c.ws, sink);
}

int ck_hmult(ck_context *ctx, int level, uint32_t *out_b, uint32_t *out_a,
             const uint32_t *b1, const uint32_t *a1, const uint32_t *b2,
             const uint32_t *a2, const uint32_t *const *key_b,
             const uint32_t *const *key_a, int rescale, void *stream) {
  CK_REQUIRE(ctx && ctx->c && out_b && out_a && b1 && a1 && b2 && a2,
             "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(valid_level(c, level), "invalid level");
  CK_REQUIRE(!rescale || moddown_plan(c, level, 2),
             "cannot rescale at this level");
  CK_REQUIRE(valid_key(c, key_b, level) && valid_key(c, key_a, level),
             "null key block");
  ImmediateSink sink(c, as_stream(stream));
  return op_hmult(c, level, out_b, out_a, b1, a1, b2, a2, key_b, key_a,
                  rescale != 0, c.ws, sink);
}

int ck_hmult_noalloc(ck_context *ctx, int level, uint32_t *out_b, uint32_t *out_a,
                     const uint32_t *b1, const uint32_t *a1, const uint32_t *b2,
                     const uint32_t *a2, const uint32_t *const *key_b,
                     const uint32_t *const *key_a, int rescale, void *stream) {
  CK_REQUIRE(ctx && ctx->c && out_b && out_a && b1 && a1 && b2 && a2,
             "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(valid_level(c, level), "invalid level");
  CK_REQUIRE(!rescale || moddown_plan(c, level, 2),
             "cannot rescale at this level");
  CK_REQUIRE(valid_key(c, key_b, level) && valid_key(c, key_a, level),
             "null key block");
  CK_REQUIRE(relin_ws_limbs(c, level) <= kScratchLimbs,
             "HMult at this level needs more workspace than the static scratch "
             "holds (192 limbs of 2^16 words)");
  CK_REQUIRE(c.n == 65536, "the static scratch is sized for N = 2^16");
  // per-kernel profiling creates HIP events, i.e. allocates
  CK_REQUIRE(!c.prof_on, "disable profiling (ck_profile_enable) for ck_hmult_noalloc");
  uint32_t *ws = nullptr;
  const int rc = scratch_ptr(&ws);
  if (rc) return rc;
  ImmediateSink sink(c, as_stream(stream));
  return op_hmult(c, level, out_b, out_a, b1, a1, b2, a2, key_b, key_a,
                  rescale != 0, ws, sink);
}

int ck_hrot(ck_context *ctx, int level, uint32_t *out_b, uint32_t *out_a,
            const uint32_t *b, const uint32_t *a, const uint32_t *const *key_b,
            const uint32_t *const *key_a, int rot, void *stream) {
  CK_REQUIRE(ctx && ctx->c && out_b && out_a && b && a, "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(valid_level(c, level), "invalid level");
  const int slots = c.n / 2;
  rot %= slots;
  if (rot < 0) rot += slots;
  const size_t bytes = (size_t)c.levels[level].num_q * kLimb * 4;
  if (rot == 0) {
    // Context::HRot copies when the rotation is a no-op
    const int rc = launch_copy(c, out_b, b, bytes, as_stream(stream));
    if (rc) return rc;
    return launch_copy(c, out_a, a, bytes, as_stream(stream));
  }
  CK_REQUIRE(valid_key(c, key_b, level) && valid_key(c, key_a, level),
             "null key block");
  ImmediateSink sink(c, as_stream(stream));
  return op_hrot_galois(c, level, out_b, out_a, b, a, key_b, key_a,
                        c.galois[rot], c.ws, sink);
}

int ck_hconj(ck_context *ctx, int level, uint32_t *out_b, uint32_t *out_a,
             const uint32_t *b, const uint32_t *a, const uint32_t *const *key_b,
             const uint32_t *const *key_a, void *stream) {
  CK_REQUIRE(ctx && ctx->c && out_b && out_a && b && a, "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(valid_level(c, level), "invalid level");
  CK_REQUIRE(valid_key(c, key_b, level) && valid_key(c, key_a, level),
             "null key block");
  ImmediateSink sink(c, as_stream(stream));
  return op_hrot_galois(c, level, out_b, out_a, b, a, key_b, key_a,
                        2u * c.n - 1, c.ws, sink);
}

int ck_rescale(ck_context *ctx, int level, int npoly, uint32_t *const *dst,
               const uint32_t *const *src, void *stream) {
  return ck_moddown(ctx, level, CK_RESCALE, npoly, dst, src, stream);
}

int ck_hrot_hoisted(ck_context *ctx, int level, int n_rot, const int *rots,
                    uint32_t *const *out_b, uint32_t *const *out_a,
                    const uint32_t *b, const uint32_t *a,
                    const uint32_t *const *key_b, const uint32_t *const *key_a,
                    int moddown, void *stream) {
  CK_REQUIRE(ctx && ctx->c && b && a && rots, "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(valid_level(c, level), "invalid level");
  CK_REQUIRE(n_rot >= 1, "n_rot must be positive");
  CK_REQUIRE(valid_ptrs(out_b, n_rot) && valid_ptrs(out_a, n_rot),
             "null output pointer");
  CK_REQUIRE(key_b && key_a, "null key table");
  for (int i = 0; i < n_rot; i++) {
    CK_REQUIRE(rots[i] >= 1 && rots[i] < c.n / 2,
               "hoisted rotation amounts must be in [1, N/2)");
    CK_REQUIRE(valid_key(c, key_b + (size_t)i * c.dnum, level) &&
                   valid_key(c, key_a + (size_t)i * c.dnum, level),
               "null key block");
    CK_REQUIRE(out_b[i] != b && out_b[i] != a && out_a[i] != b && out_a[i] != a,
               "in-place operation is not supported for hoisted rotations");
  }
  CK_REQUIRE(hoisted_ws_limbs(c, level, kMaxBsRot, moddown != 0) * kLimb <=
                 c.ws_words,
             "workspace too small");
  ImmediateSink sink(c, as_stream(stream));
  return op_hoisted_rotations(c, level, n_rot, rots, out_b, out_a, b, a, key_b,
                              key_a, moddown != 0, kMaxBsRot, c.ws, sink);
}

// Builds a plan with the context's BaseConv form and key pass, or with
// autotuning all three BaseConv forms (fused into the forward column pass,
// on the FP64 matrix cores, on the int8 matrix cores) and, when key_dim,
// both key passes (the ModUp's row stages fused with the key inner product,
// or a row pass plus a separate accumulation): times each and keeps the
// fastest. build(bconv form, key pass (-1: the context's), item chunks
// (0: the builder's default), &plan). With chunk_items > 0 (HMult plans),
// the fastest form is then also built with its items in 2, 4 and 8 chunks
// (at least two items each) and the fastest of those kept.
extern "C++" template <typename Build>
static int build_tuned(const Context &c, Build &build, bool key_dim, int chunk_items,
                       Plan **out) {
  *out = nullptr;
  if (!c.plan_autotune) return build(c.bconv_mfma < 0 ? 2 : c.bconv_mfma, -1, 0, out);
  Plan *best = nullptr;
  double best_us = -1.0;
  double us[3] = {-1.0, -1.0, -1.0};
  std::vector<Plan::Candidate> rec;
  int best_f = 0, best_kr = -1;
  auto consider = [&](int f, int kr, int chunks) -> int {
    Plan *p = nullptr;
    int rc = build(f, kr, chunks, &p);
    if (rc) return rc;
    const double t = plan_time_us(*p, 3);
    rec.push_back({f, kr < 0 ? (c.fuse_key_rows ? 1 : 0) : kr, p->chunks, t});
    if (t >= 0 && (us[f] < 0 || t < us[f])) us[f] = t;
    if (t >= 0 && (!best || t < best_us)) {
      plan_destroy(best);
      best = p;
      best_us = t;
      best_f = f;
      best_kr = kr;
    } else {
      plan_destroy(p);
    }
    return kOk;
  };
  for (int f = 0; f < 3; f++)
    for (int kr : key_dim ? std::vector<int>{1, 0} : std::vector<int>{-1}) {
      const int rc = consider(f, kr, 1);
      if (rc) {
        plan_destroy(best);
        return rc;
      }
    }
  if (best)
    for (int k : {2, 4, 8}) {
      if (chunk_items < 2 * k) break;
      const int rc = consider(best_f, best_kr, k);
      if (rc) {
        plan_destroy(best);
        return rc;
      }
    }
  if (!best) return fail(CK_ERR_HIP, "autotuning could not time any candidate plan");
  for (int f = 0; f < 3; f++) best->tuned_us[f] = us[f];
  best->tuning = rec;
  *out = best;
  return kOk;
}

// true when one of a[0..na) is one of b[0..nb) (defined below)
extern "C++" bool aliases(const void *const *a, int na, const void *const *b, int nb);

int ck_plan_hmult(ck_context *ctx, int n, const int *levels, int rescale,
                  uint32_t *const *out_b, uint32_t *const *out_a,
                  const uint32_t *const *b1, const uint32_t *const *a1,
                  const uint32_t *const *b2, const uint32_t *const *a2,
                  const uint32_t *const *key_b, const uint32_t *const *key_a,
                  ck_plan **out) {
  CK_REQUIRE(out, "out is null");
  *out = nullptr;
  CK_REQUIRE(ctx && ctx->c && levels, "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(n >= 1, "a plan needs at least one ciphertext");
  CK_REQUIRE(valid_ptrs(out_b, n) && valid_ptrs(out_a, n) && valid_ptrs(b1, n) &&
                 valid_ptrs(a1, n) && valid_ptrs(b2, n) && valid_ptrs(a2, n),
             "null pointer");
  if (c.plan_autotune) {
    // the timed candidate runs write the outputs before the caller's first
    // run: they must not overwrite the inputs
    const uint32_t *const *ins[4] = {b1, a1, b2, a2};
    for (const uint32_t *const *in : ins)
      CK_REQUIRE(!aliases((const void *const *)out_b, n, (const void *const *)in, n) &&
                     !aliases((const void *const *)out_a, n, (const void *const *)in, n),
                 "an autotuned plan's outputs must not alias its inputs");
  }
  size_t total = 0;
  for (int i = 0; i < n; i++) {
    CK_REQUIRE(valid_level(c, levels[i]), "invalid level");
    CK_REQUIRE(!rescale || moddown_plan(c, levels[i], 2),
               "cannot rescale at this level");
    CK_REQUIRE(valid_key(c, key_b, levels[i]) && valid_key(c, key_a, levels[i]),
               "null key block");
    total += relin_ws_limbs(c, levels[i]);
  }
  // default item chunks: at most 512 MiB of workspace each, so that a
  // chunk's step outputs stay in the 256 MB last-level cache for the next
  // step (64 HMults at 8..24 limbs: 4 chunks, 7-9 % faster than one)
  const size_t ws_bytes = total * kLimb * sizeof(uint32_t);
  const int default_chunks =
      (int)std::min<size_t>((size_t)n, std::max<size_t>(1, (ws_bytes + (512ull << 20) - 1) >> 29));
  auto build = [&](int form, int key_rows, int chunks, Plan **res) {
    Plan *p = new Plan;
    p->ctx = &c;
    p->items = n;
    p->bconv_form = form;
    p->bconv_mfma = form != 0;
    p->key_rows = key_rows >= 0 ? key_rows : (c.fuse_key_rows ? 1 : 0);
    p->chunks = std::max(1, std::min(chunks > 0 ? chunks : default_chunks, n));
    uint32_t *ws = nullptr;
    int rc = plan_alloc_workspace(*p, total, &ws);
    PlanSink sink(*p);
    // items i = k, k + chunks, ... form chunk k (levels mix alike); a
    // chunk's workspaces are contiguous
    size_t at = 0;
    for (int k = 0; k < p->chunks && !rc; k++) {
      sink.begin_chunk();
      for (int i = k; i < n && !rc; i += p->chunks) {
        sink.begin_item();
        rc = op_hmult(c, levels[i], out_b[i], out_a[i], b1[i], a1[i], b2[i], a2[i],
                      key_b, key_a, rescale != 0, ws + at * kLimb, sink);
        at += relin_ws_limbs(c, levels[i]);
      }
    }
    if (!rc) rc = plan_finish(*p, sink.mismatch());
    if (rc) {
      plan_destroy(p);
      p = nullptr;
    }
    *res = p;
    return rc;
  };
  Plan *p = nullptr;
  const int rc = build_tuned(c, build, true, n, &p);
  if (rc) return rc;
  *out = new ck_plan{p};
  return CK_OK;
}

// Builds every candidate, times each (one warm-up + 3 runs: the runs write
// the plan's outputs), keeps the fastest and records all times. A single
// candidate is built and returned untimed.
extern "C++" template <class Build>
static int tune_candidates(const std::vector<Plan::Candidate> &cands, Build &build,
                           Plan **out) {
  *out = nullptr;
  if (cands.size() == 1) return build(cands[0], out);
  Plan *best = nullptr;
  double best_us = -1.0;
  std::vector<Plan::Candidate> rec;
  for (Plan::Candidate cd : cands) {
    Plan *p = nullptr;
    const int rc = build(cd, &p);
    if (rc) {
      plan_destroy(best);
      return rc;
    }
    cd.us = plan_time_us(*p, 3);
    rec.push_back(cd);
    if (cd.us >= 0 && (best_us < 0 || cd.us < best_us)) {
      plan_destroy(best);
      best = p;
      best_us = cd.us;
    } else {
      plan_destroy(p);
    }
  }
  if (!best) return fail(CK_ERR_HIP, "autotuning could not time any candidate plan");
  best->tuning = rec;
  *out = best;
  return CK_OK;
}

extern "C++" bool aliases(const void *const *a, int na, const void *const *b, int nb) {
  for (int i = 0; i < na; i++)
    for (int j = 0; j < nb; j++)
      if (a[i] == b[j]) return true;
  return false;
}

int ck_plan_bconv(ck_context *ctx, const int *src_primes, int n_src,
                  const int *dst_primes, int n_dst, int n, const int32_t *const *src,
                  uint32_t *const *dst, ck_plan **out) {
  CK_REQUIRE(out, "out is null");
  *out = nullptr;
  CK_REQUIRE(ctx && ctx->c && src_primes && dst_primes && src && dst, "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(n >= 1, "a plan needs at least one polynomial");
  CK_REQUIRE(n_src >= 1 && n_src <= kMaxBconvSrc, "n_src must be 1..24");
  CK_REQUIRE(n_dst >= 1 && n_dst <= kMaxMfmaDst, "n_dst must be 1..64");
  for (int i = 0; i < n_src; i++) {
    CK_REQUIRE(src_primes[i] >= 0 && src_primes[i] < c.num_primes, "invalid source prime");
    for (int j = 0; j < i; j++)
      CK_REQUIRE(src_primes[j] != src_primes[i], "source primes must be distinct");
  }
  for (int m = 0; m < n_dst; m++)
    CK_REQUIRE(dst_primes[m] >= 0 && dst_primes[m] < c.num_primes,
               "invalid destination prime");
  for (int k = 0; k < n; k++) CK_REQUIRE(src[k] && dst[k], "null pointer");
  CK_REQUIRE(c.plan_bconv_form >= 0 ||
                 !aliases((const void *const *)src, n, (const void *const *)dst, n),
             "autotuned plans need outputs that do not alias inputs");
  std::vector<int> sp(src_primes, src_primes + n_src), dp(dst_primes, dst_primes + n_dst);
  std::vector<uint32_t> c1;
  std::vector<int32_t> table;
  std::vector<double> ftab;
  bconv_constants(c, sp, dp, {}, c1, table, ftab);
  auto build = [&](const Plan::Candidate &cd, Plan **res) {
    Plan *p = new Plan;
    p->ctx = &c;
    p->items = n;
    p->form = cd.form;
    int rc = kOk;
    if (cd.form == 0) {  // FP64 matrix cores
      p->bconv_mfma = true;
      p->bconv_form = 1;  // the FP64 matrix-core BaseConv step
      double *d_tab = nullptr;
      rc = plan_upload(*p, &d_tab, ftab);
      PlanSink sink(*p);
      for (int k = 0; k < n && !rc; k++) {
        sink.begin_item();
        BconvInst I;
        std::memset(&I, 0, sizeof(I));
        I.src = src[k];
        I.dst = dst[k];
        I.tab = d_tab;
        I.S = n_src;
        I.D = n_dst;
        I.canonical = 1;
        for (int m = 0; m < n_dst; m++) {
          I.dst_limb[m] = (uint16_t)m;
          I.dst_prime[m] = (uint8_t)dp[m];
        }
        rc = sink.bconv({I});
      }
      if (!rc) rc = plan_finish(*p, sink.mismatch());
    } else {  // int8 matrix cores (1) or vector unit (2)
      int64_t *d_planes = nullptr;
      double *d_ftab = nullptr;
      if (cd.form == 1) {
        std::vector<uint32_t> tab((size_t)n_dst * n_src), dq(n_dst);
        for (int m = 0; m < n_dst; m++) {
          dq[m] = c.primes[dp[m]];
          for (int i = 0; i < n_src; i++)
            tab[(size_t)m * n_src + i] =
                (uint32_t)ftab[2 * (m * n_src + i)] * 65536u +
                (uint32_t)ftab[2 * (m * n_src + i) + 1];
        }
        rc = plan_upload_i64(*p, &d_planes, i8::bconv_i8_planes(tab, n_src, n_dst, dq));
      } else {
        rc = plan_upload(*p, &d_ftab, ftab);
      }
      p->stages.emplace_back();
      PlanStage &s = p->stages.back();
      s.kind = cd.form == 1 ? PlanStage::kBconvI8 : PlanStage::kBconvValu;
      for (int k = 0; k < n; k++) {
        BconvI8Inst I;
        std::memset(&I, 0, sizeof(I));
        I.src = src[k];
        I.dst = dst[k];
        I.planes = d_planes;
        I.ftab = d_ftab;
        I.S = n_src;
        I.D = n_dst;
        I.canonical = 1;
        for (int m = 0; m < n_dst; m++) {
          I.dst_limb[m] = (uint16_t)m;
          I.dst_prime[m] = (uint8_t)dp[m];
        }
        s.bi.push_back(I);
      }
      if (!rc) rc = plan_finish(*p, false);
    }
    if (rc) {
      plan_destroy(p);
      p = nullptr;
    }
    *res = p;
    return rc;
  };
  std::vector<Plan::Candidate> cands;
  if (c.plan_bconv_form >= 0)
    cands.push_back({c.plan_bconv_form, 0, 0, -1.0});
  else
    for (int f = 0; f < 3; f++) cands.push_back({f, 0, 0, -1.0});
  Plan *p = nullptr;
  const int rc = tune_candidates(cands, build, &p);
  if (rc) return rc;
  *out = new ck_plan{p};
  return CK_OK;
}

int ck_plan_ntt(ck_context *ctx, int n, int num_main, int num_ter, int num_aux,
                int inverse, int montgomery_conversion, const uint32_t *const *src,
                uint32_t *const *dst, ck_plan **out) {
  CK_REQUIRE(out, "out is null");
  *out = nullptr;
  CK_REQUIRE(ctx && ctx->c && src && dst, "null argument");
  Context &c = *ctx->c;
  CK_REQUIRE(n >= 1, "a plan needs at least one polynomial");
  CK_REQUIRE(valid_np(c, num_main, num_ter, num_aux), "invalid NP");
  for (int k = 0; k < n; k++) CK_REQUIRE(src[k] && dst[k], "null pointer");
  // the int8 matrix-core form covers the forward transform; its column pass
  // writes a tiled intermediate over dst, which must therefore not overlap
  // any input: in-place plans take the vector form
  bool overlap = false;
  const size_t bytes = (size_t)(num_main + num_ter + num_aux) * kLimb * sizeof(uint32_t);
  for (int k = 0; k < n && !overlap; k++)
    for (int j = 0; j < n && !overlap; j++) {
      const char *d0 = (const char *)dst[k], *s0 = (const char *)src[j];
      overlap = d0 < s0 + bytes && s0 < d0 + bytes;
    }
  int form = inverse ? 0 : c.plan_ntt_form;
  CK_REQUIRE(form >= 0 || !overlap, "autotuned plans need outputs that do not alias inputs");
  if (overlap) form = 0;
  const int limbs = num_main + num_ter + num_aux;
  if (form != 0) {
    I8NttTabs T;
    const int rc = ntt8_tables(c, &T);
    if (rc) return rc;
  }
  auto build = [&](const Plan::Candidate &cd, Plan **res) {
    Plan *p = new Plan;
    p->ctx = &c;
    p->items = n;
    p->form = cd.form;
    int rc = kOk;
    if (cd.form == 0) {
      PlanSink sink(*p);
      for (int k = 0; k < n && !rc; k++) {
        sink.begin_item();
        rc = inverse ? op_intt(c, dst[k], src[k], num_main, num_ter, num_aux, 0,
                               montgomery_conversion != 0, sink)
                     : op_ntt(c, dst[k], src[k], num_main, num_ter, num_aux, 0,
                              montgomery_conversion != 0, sink);
      }
      if (!rc) rc = plan_finish(*p, sink.mismatch());
    } else {
      rc = ntt8_tables(c, &p->i8);
      std::vector<Ntt8Job> jobs;
      for (int k = 0; k < n; k++)
        for (int y = 0; y < limbs; y++)
          jobs.push_back(Ntt8Job{src[k] + (size_t)y * kLimb, dst[k] + (size_t)y * kLimb,
                                 global_prime(c, num_main, num_ter, y),
                                 montgomery_conversion != 0 ? 1 : 0});
      if (!rc) {
        plan_add_ntt8(*p, std::move(jobs), cd.blocks, cd.chunk);
        rc = plan_finish(*p, false);
      }
    }
    if (rc) {
      plan_destroy(p);
      p = nullptr;
    }
    *res = p;
    return rc;
  };
  std::vector<Plan::Candidate> cands;
  if (form >= 0) {
    cands.push_back({form, c.ntt8_blocks, c.ntt8_chunk, -1.0});
  } else {
    cands.push_back({0, 0, 0, -1.0});
    const int total = n * limbs;
    for (int blocks : {16, 8, 4})
      for (int chunk : {0, 512, 256, 128})
        if (chunk == 0 || chunk < total) cands.push_back({1, blocks, chunk, -1.0});
  }
  Plan *p = nullptr;
  const int rc = tune_candidates(cands, build, &p);
  if (rc) return rc;
  *out = new ck_plan{p};
  return CK_OK;
}

int ck_plan_hrot_hoisted(ck_context *ctx, int n, const int *levels,
                         const uint32_t *const *b, const uint32_t *const *a,
                         int n_rot, const int *rots,
                         const uint32_t *const *key_b,
                         const uint32_t *const *key_a, uint32_t *const *out_b,
                         uint32_t *const *out_a, int moddown, ck_plan **out) {
  CK_REQUIRE(out, "out is null");
  *out = nullptr;
  CK_REQUIRE(ctx && ctx->c && levels && rots && key_b && key_a, "null argument");
  const Context &c = *ctx->c;
  CK_REQUIRE(n >= 1 && n_rot >= 1, "a plan needs ciphertexts and rotations");
  CK_REQUIRE(valid_ptrs(b, n) && valid_ptrs(a, n), "null input pointer");
  CK_REQUIRE(valid_ptrs(out_b, n * n_rot) && valid_ptrs(out_a, n * n_rot),
             "null output pointer");
  if (c.plan_autotune) {
    // the timed candidate runs write the outputs before the caller's first
    // run: they must not overwrite the inputs
    const uint32_t *const *ins[2] = {b, a};
    for (const uint32_t *const *in : ins)
      CK_REQUIRE(!aliases((const void *const *)out_b, n * n_rot, (const void *const *)in, n) &&
                     !aliases((const void *const *)out_a, n * n_rot, (const void *const *)in, n),
                 "an autotuned plan's outputs must not alias its inputs");
  }
  size_t total = 0;
  std::vector<size_t> off(n);
  for (int i = 0; i < n; i++) {
    CK_REQUIRE(valid_level(c, levels[i]), "invalid level");
    for (int r = 0; r < n_rot; r++)
      CK_REQUIRE(valid_key(c, key_b + (size_t)r * c.dnum, levels[i]) &&
                     valid_key(c, key_a + (size_t)r * c.dnum, levels[i]),
                 "null key block");
    off[i] = total;
    total += hoisted_ws_limbs(c, levels[i], n_rot, moddown != 0);
  }
  for (int r = 0; r < n_rot; r++)
    CK_REQUIRE(rots[r] >= 1 && rots[r] < c.n / 2,
               "hoisted rotation amounts must be in [1, N/2)");
  // item chunks measured no faster here (the hoisted plan's per-item
  // workspace already exceeds the last-level cache): one chunk
  auto build = [&](int form, int key_rows, int, Plan **res) {
    Plan *p = new Plan;
    p->ctx = &c;
    p->items = n;
    p->bconv_form = form;
    p->bconv_mfma = form != 0;
    p->key_rows = key_rows;
    uint32_t *ws = nullptr;
    int rc = plan_alloc_workspace(*p, total, &ws);
    PlanSink sink(*p);
    for (int i = 0; i < n && !rc; i++) {
      sink.begin_item();
      rc = op_hoisted_rotations(c, levels[i], n_rot, rots, out_b + (size_t)i * n_rot,
                                out_a + (size_t)i * n_rot, b[i], a[i], key_b, key_a,
                                moddown != 0, n_rot, ws + off[i] * kLimb, sink);
    }
    if (!rc) rc = plan_finish(*p, sink.mismatch());
    if (rc) {
      plan_destroy(p);
      p = nullptr;
    }
    *res = p;
    return rc;
  };
  Plan *p = nullptr;
  const int rc = build_tuned(c, build, false, 0, &p);
  if (rc) return rc;
  *out = new ck_plan{p};
  return CK_OK;
}

int ck_plan_run(ck_plan *plan, void *stream) {
  CK_REQUIRE(plan && plan->p, "null plan");
  return plan_run(*plan->p, as_stream(stream));
}

int ck_plan_info(const ck_plan *plan, long long *out) {
  CK_REQUIRE(plan && plan->p && out, "null argument");
  out[0] = plan->p->items;
  out[1] = (long long)plan->p->stages.size();
  out[2] = plan_launches(*plan->p);
  out[3] = (long long)plan->p->workspace_bytes;
  out[4] = plan->p->bconv_mfma ? plan->p->bconv_form : 0;
  out[5] = (long long)(plan->p->tuned_us[0] * 1000.0);  // ns, -1000: not tuned
  out[6] = (long long)(plan->p->tuned_us[1] * 1000.0);
  out[7] = plan->p->key_rows >= 0 ? plan->p->key_rows : plan->p->form;
  return CK_OK;
}

int ck_plan_tuning(const ck_plan *plan, int max, int *forms, double *us) {
  if (!plan || !plan->p || max < 0 || (max && (!forms || !us))) {
    set_error("invalid plan tuning arguments");
    return -1;
  }
  const std::vector<Plan::Candidate> &t = plan->p->tuning;
  const int k = (int)t.size() < max ? (int)t.size() : max;
  for (int i = 0; i < k; i++) {
    forms[3 * i] = t[i].form;
    forms[3 * i + 1] = t[i].blocks;
    forms[3 * i + 2] = t[i].chunk;
    us[i] = t[i].us;
  }
  return (int)t.size();
}

int ck_plan_destroy(ck_plan *plan) {
  if (!plan) return CK_OK;
  plan_destroy(plan->p);
  delete plan;
  return CK_OK;
}

int ck_context_set_option(ck_context *ctx, int option, int value) {
  CK_REQUIRE(ctx && ctx->c, "null context");
  switch (option) {
    case CK_OPT_BCONV_MFMA:
      CK_REQUIRE(value >= -1 && value <= 2, "option value must be -1, 0, 1 or 2");
      ctx->c->bconv_mfma = value;
      return CK_OK;
    case CK_OPT_PLAN_AUTOTUNE:
      CK_REQUIRE(value == 0 || value == 1, "option value must be 0 or 1");
      ctx->c->plan_autotune = value != 0;
      return CK_OK;
    case CK_OPT_FUSED_KEY_ROWS:
      CK_REQUIRE(value == 0 || value == 1, "option value must be 0 or 1");
      ctx->c->fuse_key_rows = value != 0;
      return CK_OK;
    case CK_OPT_PLAN_NTT_FORM:
      CK_REQUIRE(value >= -1 && value <= 1, "option value must be -1, 0 or 1");
      ctx->c->plan_ntt_form = value;
      return CK_OK;
    case CK_OPT_PLAN_BCONV_FORM:
      CK_REQUIRE(value >= -1 && value <= 2, "option value must be -1, 0, 1 or 2");
      ctx->c->plan_bconv_form = value;
      return CK_OK;
    case CK_OPT_NTT8_BLOCKS:
      CK_REQUIRE(value == 1 || value == 2 || value == 4 || value == 8 || value == 16,
                 "option value must be 1, 2, 4, 8 or 16");
      ctx->c->ntt8_blocks = value;
      return CK_OK;
    case CK_OPT_NTT8_CHUNK:
      CK_REQUIRE(value >= 0, "option value must be >= 0");
      ctx->c->ntt8_chunk = value;
      return CK_OK;
    default:
      return fail(CK_ERR_INVALID_ARG, "unknown context option");
  }
}

int ck_profile_enable(ck_context *ctx, int on) {
  CK_REQUIRE(ctx && ctx->c, "null context");
  prof_reset(*ctx->c);
  ctx->c->prof_on = (on != 0);
  return CK_OK;
}

int ck_profile_report(ck_context *ctx, int max_entries, char *names,
                      double *total_ms, long long *calls) {
  if (!ctx || !ctx->c || max_entries < 0 || (max_entries && (!names || !total_ms || !calls))) {
    set_error("invalid profile report arguments");
    return -1;
  }
  std::vector<std::string> nm;
  std::vector<double> ms;
  std::vector<long> cnt;
  if (prof_report(*ctx->c, nm, ms, cnt)) return -1;
  const int n = (int)nm.size() < max_entries ? (int)nm.size() : max_entries;
  for (int i = 0; i < n; i++) {
    std::strncpy(names + 64 * i, nm[i].c_str(), 63);
    names[64 * i + 63] = 0;
    total_ms[i] = ms[i];
    calls[i] = cnt[i];
  }
  return n;
}

}  // extern "C"
