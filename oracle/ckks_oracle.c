/*
 * CPU oracle for the CKKS RNS/NTT key-switching path -- TEST INFRASTRUCTURE.
 *
 * Scalar C restatement of scale-snu/cheddar-fhe (32-bit word path). Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; it
 * is the checker, never the product.
 *
 * Parity status per function group (what pins it; tests/test_oracle_*.py):
 *   arithmetic (or_mont_*, or_mulmod, ...)  brute force with 128-bit ints,
 *       edge values 0, q-1, +-q*2^31, primes near 2^31.
 *   constants (generator, psi, qinv)        Python brute force of the
 *       smallest generator, psi^N = -1, psi^2N = 1, q*qinv = 1 mod 2^32.
 *   NTT / INTT                              direct evaluation
 *       out[k] = sum_j a_j psi^((2 brv(k)+1) j) at small N and sampled
 *       coefficients at N = 2^16, negacyclic convolution theorem against a
 *       schoolbook product, INTT(NTT(x)) = x.
 *   BaseConv                                Python big-integer CRT: with
 *       centred v_i, X = sum v_i * s^_i is the integer the output encodes.
 *   ModUp / ModDown / Rescale               Python big-integer CRT
 *       reconstruction (exact rounding identities) and ModDown(P * x) = x.
 *   Tensor / Permute / key switching        schoolbook products, Galois
 *       composition, and an encrypt -> HMult -> decrypt round trip against a
 *       big-integer negacyclic product.
 * Nothing in this file is "parity unpinned".
 */
#include "ckks_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;
typedef __int128 i128;

/* ====================================================================== */
/* Scalar arithmetic                                                       */
/* ====================================================================== */

/* primeutil::MultMod (include/common/PrimeUtils.h:57-62) */
uint32_t or_mulmod(uint32_t a, uint32_t b, uint32_t q) {
  return (uint32_t)(((uint64_t)a * b) % q);
}

/* primeutil::InvMod (PrimeUtils.h:72-84): extended Euclid; 0 on failure */
uint32_t or_invmod(uint32_t a, uint32_t q) {
  int64_t r0 = (int64_t)q, r1 = (int64_t)(a % q);
  int64_t t0 = 0, t1 = 1;
  if (r1 == 0) return 0;
  while (r1 != 0) {
    int64_t qq = r0 / r1;
    int64_t tmp = r0 - qq * r1;
    r0 = r1;
    r1 = tmp;
    tmp = t0 - qq * t1;
    t0 = t1;
    t1 = tmp;
  }
  if (r0 != 1) return 0;
  if (t0 < 0) t0 += q;
  return (uint32_t)t0;
}

/* primeutil::PowMod (PrimeUtils.h:86-125): exponent taken mod q-1, negative
 * exponents go through the inverse. */
uint32_t or_powmod(uint32_t a, int64_t e, uint32_t q) {
  a %= q;
  if (a == 0) return 0;
  int64_t phi = (int64_t)q - 1;
  e %= phi;
  if (e < 0) {
    a = or_invmod(a, q);
    e = -e;
  }
  uint64_t r = 1, b = a;
  uint64_t x = (uint64_t)e;
  while (x) {
    if (x & 1) r = (r * b) % q;
    b = (b * b) % q;
    x >>= 1;
  }
  return (uint32_t)r;
}

/* primeutil::FindGenerator (PrimeUtils.h:135-174): the SMALLEST generator.
 * Factor q-1 by trial division, then test r = 2, 3, ... */
uint32_t or_find_generator(uint32_t q) {
  uint32_t phi = q - 1, tmp = phi;
  uint32_t factors[32];
  int nf = 0;
  for (uint32_t i = 2; (uint64_t)i * i <= phi; i++) {
    if (tmp % i == 0) {
      factors[nf++] = i;
      while (tmp % i == 0) tmp /= i;
      if (tmp == 1) break;
    }
  }
  if (tmp > 1) factors[nf++] = tmp;
  for (uint32_t r = 2; r <= phi; r++) {
    int ok = 1;
    for (int f = 0; f < nf; f++) {
      if (or_powmod(r, phi / factors[f], q) == 1) {
        ok = 0;
        break;
      }
    }
    if (ok) return r;
  }
  return 0;
}

/* primeutil::FindPrimitiveMthRoot (PrimeUtils.h:184-196):
 * g^((q-1)/M) with g the smallest generator. 0 if M does not divide q-1. */
uint32_t or_primitive_root(int64_t m, uint32_t q) {
  if (((int64_t)q - 1) % m != 0) return 0;
  uint32_t g = or_find_generator(q);
  return or_powmod(g, ((int64_t)q - 1) / m, q);
}

/* primeutil::InvModBase (PrimeUtils.h:205-232): q^-1 mod 2^32 as a signed
 * word. The inverse is unique mod 2^32; computed here by Newton lifting. */
int32_t or_inv_mod_base(uint32_t q) {
  uint32_t x = q; /* correct to 3 bits for odd q */
  for (int i = 0; i < 5; i++) x *= 2u - q * x;
  return (int32_t)x;
}

/* primeutil::ToMontgomery (PrimeUtils.h:242-248): a * 2^32 mod q */
uint32_t or_to_mont(uint32_t a, uint32_t q) {
  return (uint32_t)((((uint64_t)a) << 32) % q);
}

/* basic::detail::__montgomery_reduction_lazy<uint32_t> (Basic.cuh:90-99):
 * for |a| <= q * 2^31 returns hi(a) - mulhi(lo(a) * qinv, q) in (-q, q). */
int32_t or_mont_reduce_lazy(int64_t a, uint32_t q, int32_t qinv) {
  int32_t lo = (int32_t)(uint32_t)(uint64_t)a;
  int32_t hi = (int32_t)(a >> 32);
  int32_t t = (int32_t)((uint32_t)lo * (uint32_t)qinv);
  int32_t m = (int32_t)(((int64_t)t * (int64_t)(int32_t)q) >> 32);
  return hi - m;
}

/* basic::ReduceMontgomery (Basic.cuh:148-155): lazy result lifted to [0,q) */
uint32_t or_mont_reduce(int64_t a, uint32_t q, int32_t qinv) {
  int32_t r = or_mont_reduce_lazy(a, q, qinv);
  if (r < 0) r += (int32_t)q;
  return (uint32_t)r;
}

/* basic::MultMontgomery (Basic.cuh:251-260): a * b * 2^-32 mod q, a,b<q */
uint32_t or_mult_mont(uint32_t a, uint32_t b, uint32_t q, int32_t qinv) {
  int64_t prod = (int64_t)(int32_t)a * (int64_t)(int32_t)b;
  return or_mont_reduce(prod, q, qinv);
}

/* basic::Normalize (Basic.cuh:226-239): [0,q) -> [-(q-1)/2, (q-1)/2] */
int32_t or_normalize(uint32_t a, uint32_t q) {
  int32_t r = (int32_t)a;
  if (a > (q >> 1)) r -= (int32_t)q;
  return r;
}

/* BitReverseInt (include/common/CommonUtils.h) over `bits` low bits */
uint32_t or_bitrev(uint32_t i, int bits) {
  uint32_t r = 0;
  for (int b = 0; b < bits; b++) {
    r = (r << 1) | (i & 1u);
    i >>= 1;
  }
  return r;
}

static inline uint32_t add_mod(uint32_t a, uint32_t b, uint32_t q) {
  uint32_t r = a + b;
  return r >= q ? r - q : r;
}
static inline uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t q) {
  return a >= b ? a - b : a + q - b;
}
static inline uint32_t reduce_i64(int64_t v, uint32_t q) {
  int64_t r = v % (int64_t)q;
  if (r < 0) r += q;
  return (uint32_t)r;
}

/* ====================================================================== */
/* Parameters                                                              */
/* ====================================================================== */

struct or_param {
  int log_n, n;
  int max_main, max_ter, alpha, L, dnum, max_level, num_primes;
  int *lv_main, *lv_ter; /* per level */
  uint32_t *q;           /* num_primes, [ter reversed | main | aux] */
  int32_t *qinv;
  uint32_t *psi;
  uint32_t *psi_rev_mont;     /* num_primes * n */
  uint32_t *psi_inv_rev_mont; /* num_primes * n */
  uint32_t *n_inv, *n_inv_mont;
  uint32_t *r_mod;     /* 2^32 mod q */
  uint32_t *r_inv;     /* 2^-32 mod q */
  uint32_t *p_prod_mont; /* L entries */
  uint32_t *galois;      /* n/2 + 1 */
};

static int g_num_threads = 1;

void or_set_num_threads(int n) {
  if (n < 1) n = 1;
  g_num_threads = n;
}
int or_get_num_threads(void) { return g_num_threads; }

/* CompareNPPair (Parameter.cu:10-15): strictly more total primes, or equal
 * total and more main primes. */
static int np_less(int am, int at, int bm, int bt) {
  int a = am + at, b = bm + bt;
  return (a < b) || (a == b && am < bm);
}

/* Parameter<word>::Parameter (Parameter.cu:20-193) minus the scale maths */
or_param *or_param_create(int log_n, const uint32_t *main_primes, int n_main,
                          const uint32_t *ter_primes, int n_ter,
                          const uint32_t *aux_primes, int n_aux,
                          const int *level_cfg, int n_levels) {
  if (log_n < 1 || log_n > 17 || n_main < 0 || n_ter < 0 || n_aux <= 0 ||
      n_levels <= 0 || n_main + n_ter <= 0)
    return NULL;
  int n = 1 << log_n;
  /* Parameter.cu:43-64: primes below 2^31 and = 1 mod 2N */
  int np = n_main + n_ter + n_aux;
  for (int i = 0; i < np; i++) {
    uint32_t pr = i < n_main ? main_primes[i]
                  : i < n_main + n_ter ? ter_primes[i - n_main]
                                       : aux_primes[i - n_main - n_ter];
    if (pr >= (1u << 31) || pr % (uint32_t)(2 * n) != 1) return NULL;
  }
  /* Parameter.cu:66-91: level configuration ordering */
  int pm = 0, pt = 0;
  for (int i = 0; i < n_levels - 1; i++) {
    int m = level_cfg[2 * i], t = level_cfg[2 * i + 1];
    if (m < 0 || m > n_main || t < 0 || t > n_ter) return NULL;
    if (!np_less(pm, pt, m, t)) return NULL;
    pm = m;
    pt = t;
  }
  if (level_cfg[2 * (n_levels - 1)] != n_main ||
      level_cfg[2 * (n_levels - 1) + 1] != n_ter)
    return NULL;

  or_param *p = (or_param *)calloc(1, sizeof(or_param));
  p->log_n = log_n;
  p->n = n;
  p->max_main = n_main;
  p->max_ter = n_ter;
  p->alpha = n_aux;
  p->L = n_main + n_ter;
  p->dnum = (p->L + n_aux - 1) / n_aux; /* Parameter.cu:29 */
  p->max_level = n_levels - 1;
  p->num_primes = np;
  p->lv_main = (int *)malloc(sizeof(int) * n_levels);
  p->lv_ter = (int *)malloc(sizeof(int) * n_levels);
  for (int i = 0; i < n_levels; i++) {
    p->lv_main[i] = level_cfg[2 * i];
    p->lv_ter[i] = level_cfg[2 * i + 1];
  }
  /* Parameter.cu:103-111: flattened order [ter reversed | main], then aux */
  p->q = (uint32_t *)malloc(sizeof(uint32_t) * np);
  for (int i = 0; i < n_ter; i++) p->q[i] = ter_primes[n_ter - 1 - i];
  for (int i = 0; i < n_main; i++) p->q[n_ter + i] = main_primes[i];
  for (int i = 0; i < n_aux; i++) p->q[p->L + i] = aux_primes[i];

  p->qinv = (int32_t *)malloc(sizeof(int32_t) * np);
  p->psi = (uint32_t *)malloc(sizeof(uint32_t) * np);
  p->n_inv = (uint32_t *)malloc(sizeof(uint32_t) * np);
  p->n_inv_mont = (uint32_t *)malloc(sizeof(uint32_t) * np);
  p->r_mod = (uint32_t *)malloc(sizeof(uint32_t) * np);
  p->r_inv = (uint32_t *)malloc(sizeof(uint32_t) * np);
  p->psi_rev_mont = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)np * n);
  p->psi_inv_rev_mont = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)np * n);

  /* NTTHandler::PopulateTwiddleFactors (src/core/NTT.cu:1233-1299) */
  for (int i = 0; i < np; i++) {
    uint32_t q = p->q[i];
    p->qinv[i] = or_inv_mod_base(q);
    uint32_t psi = or_primitive_root(2 * (int64_t)n, q);
    uint32_t psi_inv = or_invmod(psi, q);
    p->psi[i] = psi;
    p->n_inv[i] = or_invmod((uint32_t)n, q);
    p->n_inv_mont[i] = or_to_mont(p->n_inv[i], q);
    p->r_mod[i] = or_to_mont(1, q);
    p->r_inv[i] = or_invmod(p->r_mod[i], q);
    uint32_t *fw = p->psi_rev_mont + (size_t)i * n;
    uint32_t *iv = p->psi_inv_rev_mont + (size_t)i * n;
    uint64_t pw = 1, pwi = 1;
    for (int j = 0; j < n; j++) {
      uint32_t r = or_bitrev((uint32_t)j, log_n);
      fw[r] = or_to_mont((uint32_t)pw, q);
      iv[r] = or_to_mont((uint32_t)pwi, q);
      pw = (pw * psi) % q;
      pwi = (pwi * psi_inv) % q;
    }
  }
  /* Context ctor (src/core/Context.cpp:85-98): P mod q_i, Montgomery form */
  p->p_prod_mont = (uint32_t *)malloc(sizeof(uint32_t) * p->L);
  for (int i = 0; i < p->L; i++) {
    uint32_t v = 1;
    for (int j = 0; j < n_aux; j++) v = or_mulmod(v, p->q[p->L + j], p->q[i]);
    p->p_prod_mont[i] = or_to_mont(v, p->q[i]);
  }
  /* Parameter.cu:93-101: galois factors 5^i mod 2N, slot count N/2 */
  int slots = n / 2;
  p->galois = (uint32_t *)malloc(sizeof(uint32_t) * (slots + 1));
  p->galois[0] = 1;
  p->galois[slots] = 1;
  for (int i = 1; i < slots; i++)
    p->galois[i] = (uint32_t)(((uint64_t)p->galois[i - 1] * 5u) % (2u * n));
  return p;
}

void or_param_destroy(or_param *p) {
  if (!p) return;
  free(p->lv_main);
  free(p->lv_ter);
  free(p->q);
  free(p->qinv);
  free(p->psi);
  free(p->n_inv);
  free(p->n_inv_mont);
  free(p->r_mod);
  free(p->r_inv);
  free(p->psi_rev_mont);
  free(p->psi_inv_rev_mont);
  free(p->p_prod_mont);
  free(p->galois);
  free(p);
}

void or_param_info(const or_param *p, int *out) {
  out[0] = p->log_n;
  out[1] = p->n;
  out[2] = p->max_main;
  out[3] = p->max_ter;
  out[4] = p->alpha;
  out[5] = p->L;
  out[6] = p->dnum;
  out[7] = p->max_level;
  out[8] = p->num_primes;
  out[9] = 0;
}

uint32_t or_prime(const or_param *p, int pidx) { return p->q[pidx]; }
int32_t or_prime_qinv(const or_param *p, int pidx) { return p->qinv[pidx]; }
uint32_t or_prime_psi(const or_param *p, int pidx) { return p->psi[pidx]; }
const uint32_t *or_psi_rev_mont(const or_param *p, int pidx) {
  return p->psi_rev_mont + (size_t)pidx * p->n;
}
const uint32_t *or_psi_inv_rev_mont(const or_param *p, int pidx) {
  return p->psi_inv_rev_mont + (size_t)pidx * p->n;
}
uint32_t or_p_prod_mont(const or_param *p, int pidx) {
  return p->p_prod_mont[pidx];
}

/* Parameter::LevelToNP (Parameter.cu:287-296), levels >= 0 only */
int or_level_np(const or_param *p, int level, int *num_main, int *num_ter) {
  if (level < 0 || level > p->max_level) return -1;
  *num_main = p->lv_main[level];
  *num_ter = p->lv_ter[level];
  return 0;
}

/* Parameter::NPToLevel (Parameter.cu:298-315) */
int or_np_level(const or_param *p, int num_main, int num_ter) {
  for (int l = 0; l <= p->max_level; l++)
    if (p->lv_main[l] == num_main && p->lv_ter[l] == num_ter) return l;
  return -1;
}

/* Parameter::GetPrimeVector (Parameter.cu:317-325) as a global index */
int or_np_prime(const or_param *p, int num_main, int num_ter, int limb) {
  int num_q = num_main + num_ter;
  if (limb < num_q) return p->max_ter - num_ter + limb;
  return p->L + (limb - num_q);
}

/* ElementWiseHandler::PermuteAmountToGaloisFactor (ElementWise.cu:955-964) */
uint32_t or_galois_factor(const or_param *p, int r) {
  if (r == -1) return 2u * p->n - 1;
  if (r < 0 || r > p->n / 2) return 0;
  return p->galois[r];
}

/* ====================================================================== */
/* RNG                                                                     */
/* ====================================================================== */

uint64_t or_splitmix64(uint64_t *state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

typedef struct {
  uint64_t s[4];
} xoshiro;

static inline uint64_t rotl(uint64_t x, int k) {
  return (x << k) | (x >> (64 - k));
}
static void xo_seed(xoshiro *x, uint64_t seed) {
  uint64_t st = seed;
  for (int i = 0; i < 4; i++) x->s[i] = or_splitmix64(&st);
}
static inline uint64_t xo_next(xoshiro *x) {
  uint64_t *s = x->s;
  uint64_t result = rotl(s[1] * 5, 7) * 9;
  uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}
static inline uint32_t xo_uniform(xoshiro *x, uint32_t q) {
  int bits = 32 - __builtin_clz(q);
  for (;;) {
    uint32_t r = (uint32_t)(xo_next(x) >> (64 - bits));
    if (r < q) return r;
  }
}

void or_fill_uniform_primes(const or_param *p, const int *pidx, int nlimbs,
                            uint64_t seed, uint32_t *out) {
  xoshiro x;
  xo_seed(&x, seed);
  for (int t = 0; t < nlimbs; t++) {
    uint32_t q = p->q[pidx[t]];
    uint32_t *o = out + (size_t)t * p->n;
    for (int j = 0; j < p->n; j++) o[j] = xo_uniform(&x, q);
  }
}

void or_fill_uniform(const or_param *p, int num_main, int num_ter,
                     int num_aux, uint64_t seed, uint32_t *out) {
  int nl = num_main + num_ter + num_aux;
  int *idx = (int *)malloc(sizeof(int) * nl);
  for (int t = 0; t < nl; t++) idx[t] = or_np_prime(p, num_main, num_ter, t);
  or_fill_uniform_primes(p, idx, nl, seed, out);
  free(idx);
}

/* ====================================================================== */
/* NTT                                                                     */
/* ====================================================================== */

/* Forward negacyclic NTT, Cooley-Tukey with bit-reversed psi powers
 * (Longa-Naehrig), as NTTPhase1 + NTTPhase2 compute it
 * (src/core/NTT.cu:231-454, butterfly NTTUtils.cuh:19-34): stage m uses the
 * twiddle psi_rev[m + i] for block i; output position k holds
 * a(psi^(2 brv(k) + 1)). Twiddles are stored x*2^32 and multiplied with a
 * Montgomery product, so the transform preserves the input's form. */
void or_ntt_limb(const or_param *p, int pidx, uint32_t *a, int mont_conv) {
  const int n = p->n;
  const uint32_t q = p->q[pidx];
  const int32_t qinv = p->qinv[pidx];
  const uint32_t *w = p->psi_rev_mont + (size_t)pidx * n;
  if (mont_conv) {
    /* NTT(..., montgomery_conversion=true): x * (R^2) * R^-1 (NTT.cu:
     * 628-640 with montgomery_converter_ = ToMont(ToMont(1))) */
    uint32_t r2 = or_to_mont(or_to_mont(1, q), q);
    for (int j = 0; j < n; j++) a[j] = or_mult_mont(a[j], r2, q, qinv);
  }
  for (int m = 1, t = n >> 1; m < n; m <<= 1, t >>= 1) {
    for (int i = 0; i < m; i++) {
      const uint32_t s = w[m + i];
      uint32_t *x = a + 2 * i * t;
      for (int j = 0; j < t; j++) {
        uint32_t u = x[j];
        uint32_t v = or_mult_mont(x[j + t], s, q, qinv);
        x[j] = add_mod(u, v, q);
        x[j + t] = sub_mod(u, v, q);
      }
    }
  }
}

/* Inverse transform, Gentleman-Sande with bit-reversed psi^-1 powers
 * (INTTPhase1 + INTTPhase2, NTT.cu:23-229, ButterflyINTT NTTUtils.cuh:
 * 36-58): a' = a + b, b' = (a - b) * w. No N^-1 factor. */
void or_intt_raw_limb(const or_param *p, int pidx, uint32_t *a) {
  const int n = p->n;
  const uint32_t q = p->q[pidx];
  const int32_t qinv = p->qinv[pidx];
  const uint32_t *w = p->psi_inv_rev_mont + (size_t)pidx * n;
  for (int m = n >> 1, t = 1; m >= 1; m >>= 1, t <<= 1) {
    for (int i = 0; i < m; i++) {
      const uint32_t s = w[m + i];
      uint32_t *x = a + 2 * i * t;
      for (int j = 0; j < t; j++) {
        uint32_t u = x[j], v = x[j + t];
        x[j] = add_mod(u, v, q);
        x[j + t] = or_mult_mont(sub_mod(u, v, q), s, q, qinv);
      }
    }
  }
}

void or_ntt(const or_param *p, int num_main, int num_ter, int num_aux,
            uint32_t *data, int mont_conv) {
  int nl = num_main + num_ter + num_aux;
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_num_threads)
  for (int t = 0; t < nl; t++)
    or_ntt_limb(p, or_np_prime(p, num_main, num_ter, t),
                data + (size_t)t * p->n, mont_conv);
}

/* MultConst / MultConstNormalize elem funcs (NTTUtils.cuh:360-378) */
static inline int32_t mult_const_out(uint32_t x, uint32_t c, uint32_t q,
                                     int32_t qinv, int normalize) {
  uint32_t r = or_mult_mont(x, c, q, qinv);
  return normalize ? or_normalize(r, q) : (int32_t)r;
}

void or_intt_mult_const(const or_param *p, const int *pidx, int nlimbs,
                        const uint32_t *src, const uint32_t *consts,
                        int normalize, int32_t *out) {
  const int n = p->n;
#pragma omp parallel num_threads(g_num_threads)
  {
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * n);
#pragma omp for schedule(dynamic, 1)
    for (int t = 0; t < nlimbs; t++) {
      int pi = pidx[t];
      memcpy(tmp, src + (size_t)t * n, sizeof(uint32_t) * n);
      or_intt_raw_limb(p, pi, tmp);
      for (int j = 0; j < n; j++)
        out[(size_t)t * n + j] =
            mult_const_out(tmp[j], consts[t], p->q[pi], p->qinv[pi], normalize);
    }
    free(tmp);
  }
}

/* NTTHandler::INTT (NTT.cu:673-741): constant N^-1 (leave Montgomery form)
 * or N^-1 * R (keep it), canonical output. */
void or_intt(const or_param *p, int num_main, int num_ter, int num_aux,
             uint32_t *data, int mont_conv) {
  int nl = num_main + num_ter + num_aux;
  int *idx = (int *)malloc(sizeof(int) * nl);
  uint32_t *c = (uint32_t *)malloc(sizeof(uint32_t) * nl);
  for (int t = 0; t < nl; t++) {
    idx[t] = or_np_prime(p, num_main, num_ter, t);
    c[t] = mont_conv ? p->n_inv[idx[t]] : p->n_inv_mont[idx[t]];
  }
  or_intt_mult_const(p, idx, nl, data, c, 0, (int32_t *)data);
  free(idx);
  free(c);
}

/* ====================================================================== */
/* Base conversion                                                          */
/* ====================================================================== */

/* ModSwitchHandler::PopulateModSwitchConstants (ModSwitch.cu:320-366) with
 * kFuseModDownEpilogue = true (Parameter.h:13-20). */
void or_bconv_tables(const or_param *p, const int *src_pidx, int s,
                     const int *dst_pidx, int d, const int *restore_pidx,
                     int n_restore, uint32_t *const1, int32_t *table) {
  for (int i = 0; i < s; i++) {
    uint32_t mod = p->q[src_pidx[i]];
    uint32_t v = 1;
    for (int j = 0; j < s; j++)
      if (j != i) v = or_mulmod(v, p->q[src_pidx[j]], mod);
    v = or_mulmod(v, (uint32_t)p->n, mod); /* inverse N of the INTT */
    v = or_invmod(v, mod);
    for (int r = 0; r < n_restore; r++)
      v = or_mulmod(v, p->q[restore_pidx[r]], mod);
    const1[i] = v; /* deliberately not in Montgomery form */
  }
  for (int m = 0; m < d; m++) {
    uint32_t mod = p->q[dst_pidx[m]];
    for (int i = 0; i < s; i++) {
      uint32_t acc = 1;
      for (int j = 0; j < s; j++)
        if (j != i) acc = or_mulmod(acc, p->q[src_pidx[j]], mod);
      acc = or_to_mont(or_to_mont(acc, mod), mod); /* double Montgomery */
      table[m * s + i] = or_normalize(acc, mod);
    }
  }
}

/* ModSwitchMatrixMult (ModSwitch.cu:23-171): y = ReduceMontgomery(sum_i
 * v_i * T_mi). The kernel's lazy int64 accumulation with re-centring by
 * q*2^32 does not change the value mod q, so the exact sum is used here. */
void or_bconv(const or_param *p, const int *src_pidx, int s,
              const int32_t *src, const int *dst_pidx, int d,
              const int32_t *table, uint32_t *dst) {
  (void)src_pidx;
  const int n = p->n;
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_num_threads)
  for (int m = 0; m < d; m++) {
    uint32_t q = p->q[dst_pidx[m]];
    int32_t qinv = p->qinv[dst_pidx[m]];
    const int32_t *trow = table + (size_t)m * s;
    uint32_t *o = dst + (size_t)m * n;
    for (int k = 0; k < n; k++) {
      i128 acc = 0;
      for (int i = 0; i < s; i++)
        acc += (i128)((int64_t)src[(size_t)i * n + k] * (int64_t)trow[i]);
      int64_t r = (int64_t)(acc % (i128)q);
      if (r < 0) r += q;
      o[k] = or_mont_reduce(r, q, qinv);
    }
  }
}

/* ====================================================================== */
/* ModUp                                                                    */
/* ====================================================================== */

/* beta_ of ModSwitchHandler (ModSwitch.cu:181-185) */
int or_modup_beta(const or_param *p, int level) {
  return (p->lv_main[level] + p->max_ter + p->alpha - 1) / p->alpha;
}

static void level_prime_list(const or_param *p, int nm, int nt, int na,
                             int *out) {
  for (int t = 0; t < nm + nt + na; t++) out[t] = or_np_prime(p, nm, nt, t);
}

/* ModSwitchHandler::ModUp (ModSwitch.cu:429-509) with the constants of the
 * constructor (ModSwitch.cu:196-231). */
void or_modup(const or_param *p, int level, const uint32_t *src,
              uint32_t *out) {
  const int n = p->n, alpha = p->alpha;
  int nm = p->lv_main[level], nt = p->lv_ter[level];
  int num_q = nm + nt, tot = num_q + alpha;
  int pad = p->max_ter - nt;
  int padded = num_q + pad;
  int beta = or_modup_beta(p, level);
  int *lp = (int *)malloc(sizeof(int) * tot);
  level_prime_list(p, nm, nt, alpha, lp);
  memset(out, 0, sizeof(uint32_t) * (size_t)beta * tot * n);

  int32_t *v = (int32_t *)malloc(sizeof(int32_t) * (size_t)num_q * n);
  uint32_t *c1 = (uint32_t *)malloc(sizeof(uint32_t) * num_q);
  int *dst_idx = (int *)malloc(sizeof(int) * tot);
  int32_t *table = (int32_t *)malloc(sizeof(int32_t) * (size_t)tot * alpha);
  uint32_t *gc1 = (uint32_t *)malloc(sizeof(uint32_t) * alpha);
  uint32_t *y = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)tot * n);

  /* mod_up1_: per q limb, const1 of its group */
  for (int g = 0; g < beta; g++) {
    int s0 = g * alpha, s1 = (g + 1) * alpha < padded ? (g + 1) * alpha
                                                      : padded;
    if (s1 <= pad) continue;
    s0 = s0 - pad > 0 ? s0 - pad : 0;
    s1 -= pad;
    int d = 0;
    for (int t = 0; t < s0; t++) dst_idx[d++] = lp[t];
    for (int t = s1; t < tot; t++) dst_idx[d++] = lp[t];
    or_bconv_tables(p, lp + s0, s1 - s0, dst_idx, d, NULL, 0, gc1, table);
    for (int t = s0; t < s1; t++) c1[t] = gc1[t - s0];
  }
  /* INTTAndMultConst(normalize = true) on all q limbs */
  or_intt_mult_const(p, lp, num_q, src, c1, 1, v);

  for (int g = 0; g < beta; g++) {
    int s0 = g * alpha, s1 = (g + 1) * alpha < padded ? (g + 1) * alpha
                                                      : padded;
    if (s1 <= pad) continue;
    s0 = s0 - pad > 0 ? s0 - pad : 0;
    s1 -= pad;
    int s = s1 - s0;
    uint32_t *og = out + (size_t)g * tot * n;
    /* own limbs are copied from the NTT-domain source (ModSwitch.cu:466) */
    memcpy(og + (size_t)s0 * n, src + (size_t)s0 * n,
           sizeof(uint32_t) * (size_t)s * n);
    int d = 0;
    int *dpos = (int *)malloc(sizeof(int) * tot);
    for (int t = 0; t < s0; t++) {
      dpos[d] = t;
      dst_idx[d++] = lp[t];
    }
    for (int t = s1; t < tot; t++) {
      dpos[d] = t;
      dst_idx[d++] = lp[t];
    }
    or_bconv_tables(p, lp + s0, s, dst_idx, d, NULL, 0, gc1, table);
    or_bconv(p, lp + s0, s, v + (size_t)s0 * n, dst_idx, d, table, y);
    /* NTTForModUp: forward NTT without Montgomery conversion
     * (kFuseMontgomery: BaseConv already returned x * R) */
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_num_threads)
    for (int k = 0; k < d; k++) {
      or_ntt_limb(p, dst_idx[k], y + (size_t)k * n, 0);
      memcpy(og + (size_t)dpos[k] * n, y + (size_t)k * n,
             sizeof(uint32_t) * n);
    }
    free(dpos);
  }
  free(lp);
  free(v);
  free(c1);
  free(dst_idx);
  free(table);
  free(gc1);
  free(y);
}

/* ====================================================================== */
/* ModDown / Rescale / ModDownAndRescale                                    */
/* ====================================================================== */

/* ModSwitchHandler ctor (ModSwitch.cu:233-303) + ModDownWorker
 * (ModSwitch.cu:566-684) + NTTPhase2ForModDown epilogue (NTT.cu:456-601). */
int or_moddown(const or_param *p, int level, int kind, const uint32_t *src,
               uint32_t *dst) {
  const int n = p->n, alpha = p->alpha;
  if (level < 0 || level > p->max_level) return -1;
  if (kind != OR_MODDOWN && level == 0) return -1;
  int nm = p->lv_main[level], nt = p->lv_ter[level];
  int num_q = nm + nt;
  int num_src_aux = (kind == OR_RESCALE) ? 0 : alpha;
  int dm = nm, dt = nt; /* np_dst */
  if (kind != OR_MODDOWN) {
    dm = p->lv_main[level - 1];
    dt = p->lv_ter[level - 1];
  }
  int next_num_q = dm + dt;
  int im = nm < dm ? nm : dm, it = nt < dt ? nt : dt; /* np_non_intt */
  int src_total = num_q + num_src_aux;
  int src_len = src_total - (im + it);
  int dst_len = dm + dt;

  /* Source primes (in the order BaseConv sees them) and restore range */
  int *src_p = (int *)malloc(sizeof(int) * (src_len > 0 ? src_len : 1));
  int *src_limb = (int *)malloc(sizeof(int) * (src_len > 0 ? src_len : 1));
  int *dst_p = (int *)malloc(sizeof(int) * dst_len);
  int restore_start = 0, restore_end = 0;
  int pad_start = 0, pad_end = dst_len;
  level_prime_list(p, dm, dt, 0, dst_p);
  int ns = 0;
  if (kind == OR_MODDOWN) {
    /* ModDown constants: P -> Q (ModSwitch.cu:233-243) */
  } else {
    int main_diff = nm - dm, ter_diff = nt - dt;
    if (ter_diff > 0) {
      if (main_diff > 0) goto fail;
      for (int t = 0; t < ter_diff; t++) {
        src_limb[ns] = t;
        src_p[ns++] = or_np_prime(p, nm, nt, t);
      }
      pad_start = 0;
      pad_end = num_q - ter_diff;
      restore_start = next_num_q + main_diff;
      restore_end = next_num_q;
    } else if (main_diff > 0) {
      for (int t = num_q - main_diff; t < num_q; t++) {
        src_limb[ns] = t;
        src_p[ns++] = or_np_prime(p, nm, nt, t);
      }
      pad_start = -ter_diff;
      pad_end = num_q - main_diff - ter_diff;
      restore_start = 0;
      restore_end = -ter_diff;
    } else {
      goto fail;
    }
  }
  for (int a = 0; a < num_src_aux; a++) {
    src_limb[ns] = num_q + a;
    src_p[ns++] = p->L + a;
  }
  if (ns != src_len) goto fail;
  (void)pad_start;
  (void)pad_end;
  {
    int n_restore = restore_end - restore_start;
    int *restore_p = dst_p + restore_start;
    uint32_t *c1 = (uint32_t *)malloc(sizeof(uint32_t) * src_len);
    int32_t *table =
        (int32_t *)malloc(sizeof(int32_t) * (size_t)src_len * dst_len);
    or_bconv_tables(p, src_p, src_len, dst_p, dst_len, restore_p, n_restore,
                    c1, table);
    /* PopulateModDownEpilogueConstants (ModSwitch.cu:368-411) */
    uint32_t *inv_prod = (uint32_t *)malloc(sizeof(uint32_t) * dst_len);
    for (int i = 0; i < dst_len; i++) {
      uint32_t mod = p->q[dst_p[i]];
      uint32_t v = 1;
      for (int j = 0; j < src_len; j++) v = or_mulmod(v, p->q[src_p[j]], mod);
      inv_prod[i] = or_to_mont(or_invmod(v, mod), mod);
    }
    int use_padding = (kind != OR_MODDOWN);
    uint32_t *padding = (uint32_t *)malloc(sizeof(uint32_t) * dst_len);
    for (int i = 0; i < dst_len - n_restore; i++) {
      int mi = i >= restore_start ? i + n_restore : i;
      uint32_t mod = p->q[dst_p[mi]];
      uint32_t v = 1;
      for (int j = restore_start; j < restore_end; j++)
        v = or_mulmod(v, p->q[dst_p[j]], mod);
      padding[i] = or_to_mont(v, mod);
    }

    /* INTTForModDown: INTT of the removed limbs, x const1, centred */
    uint32_t *s_in = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)src_len * n);
    for (int k = 0; k < src_len; k++)
      memcpy(s_in + (size_t)k * n, src + (size_t)src_limb[k] * n,
             sizeof(uint32_t) * n);
    int32_t *v = (int32_t *)malloc(sizeof(int32_t) * (size_t)src_len * n);
    or_intt_mult_const(p, src_p, src_len, s_in, c1, 1, v);
    /* BaseConv to the destination primes, then NTT */
    uint32_t *y = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)dst_len * n);
    or_bconv(p, src_p, src_len, v, dst_p, dst_len, table, y);

    int src2_start = dt - it;
    int src2_end = src2_start + (im + it);
    int src2_offset = nt - dt > 0 ? nt - dt : 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_num_threads)
    for (int yi = 0; yi < dst_len; yi++) {
      int pi = dst_p[yi];
      uint32_t q = p->q[pi];
      int32_t qinv = p->qinv[pi];
      uint32_t *yl = y + (size_t)yi * n;
      or_ntt_limb(p, pi, yl, 0);
      int idx2 = yi - src2_start;
      int exists = (idx2 >= 0 && yi < src2_end);
      const uint32_t *s2 = src + (size_t)(src2_offset + idx2) * n;
      for (int k = 0; k < n; k++) {
        uint32_t r = 0;
        if (exists) {
          r = s2[k];
          if (use_padding) r = or_mult_mont(r, padding[idx2], q, qinv);
        }
        r = sub_mod(r, yl[k], q);
        dst[(size_t)yi * n + k] = or_mult_mont(r, inv_prod[yi], q, qinv);
      }
    }
    free(c1);
    free(table);
    free(inv_prod);
    free(padding);
    free(s_in);
    free(v);
    free(y);
  }
  free(src_p);
  free(src_limb);
  free(dst_p);
  return 0;
fail:
  free(src_p);
  free(src_limb);
  free(dst_p);
  return -1;
}

/* ====================================================================== */
/* Element-wise                                                             */
/* ====================================================================== */

#define EW_LOOP(body)                                                     \
  do {                                                                    \
    const int n = p->n;                                                   \
    _Pragma("omp parallel for schedule(static) num_threads(g_num_threads)") \
    for (int t = 0; t < nlimbs; t++) {                                    \
      const uint32_t q = p->q[pidx[t]];                                   \
      const int32_t qinv = p->qinv[pidx[t]];                              \
      (void)qinv;                                                         \
      for (size_t k = (size_t)t * n; k < (size_t)(t + 1) * n; k++) {      \
        body;                                                             \
      }                                                                   \
    }                                                                     \
  } while (0)

/* Sum / Sub / Neg / Mult kernels (ElementWise.cu:32-143) */
void or_add(const or_param *p, const int *pidx, int nlimbs,
            const uint32_t *x, const uint32_t *y, uint32_t *out) {
  EW_LOOP(out[k] = add_mod(x[k], y[k], q));
}
void or_sub(const or_param *p, const int *pidx, int nlimbs,
            const uint32_t *x, const uint32_t *y, uint32_t *out) {
  EW_LOOP(out[k] = sub_mod(x[k], y[k], q));
}
void or_neg(const or_param *p, const int *pidx, int nlimbs,
            const uint32_t *x, uint32_t *out) {
  EW_LOOP(out[k] = x[k] ? q - x[k] : 0);
}
void or_mult(const or_param *p, const int *pidx, int nlimbs,
             const uint32_t *x, const uint32_t *y, uint32_t *out) {
  EW_LOOP(out[k] = or_mult_mont(x[k], y[k], q, qinv));
}
/* CPAccum<1, const> = MultConst (ElementWise.cu:225-287) */
void or_mult_const(const or_param *p, const int *pidx, int nlimbs,
                   const uint32_t *x, const uint32_t *consts, uint32_t *out) {
  EW_LOOP(out[k] = or_mult_mont(x[k], consts[t], q, qinv));
}

/* Tensor (ElementWise.cu:504-546): Karatsuba middle term */
void or_tensor(const or_param *p, const int *pidx, int nlimbs,
               const uint32_t *b1, const uint32_t *a1, const uint32_t *b2,
               const uint32_t *a2, uint32_t *d0, uint32_t *d1, uint32_t *d2) {
  EW_LOOP({
    uint32_t bb = or_mult_mont(b1[k], b2[k], q, qinv);
    uint32_t aa = or_mult_mont(a1[k], a2[k], q, qinv);
    uint32_t mid = or_mult_mont(add_mod(b1[k], a1[k], q),
                                add_mod(b2[k], a2[k], q), q, qinv);
    d0[k] = bb;
    d1[k] = sub_mod(sub_mod(mid, bb, q), aa, q);
    d2[k] = aa;
  });
}

/* Permute (ElementWise.cu:370-391): NTT-domain automorphism X -> X^g,
 * dst[i] = src[brv_{logN+1}((brv_{logN+1}(i) + 1) * g - 1)] */
void or_permute(const or_param *p, int nlimbs, uint32_t galois_factor,
                const uint32_t *src, uint32_t *dst) {
  const int n = p->n, lb = p->log_n + 1;
  const uint32_t mask = (2u << p->log_n) - 1;
#pragma omp parallel for schedule(static) num_threads(g_num_threads)
  for (int t = 0; t < nlimbs; t++) {
    for (int i = 0; i < n; i++) {
      uint32_t e = or_bitrev((uint32_t)i, lb) + 1;
      uint32_t s = or_bitrev(((e * galois_factor) - 1) & mask, lb);
      dst[(size_t)t * n + i] = src[(size_t)t * n + s];
    }
  }
}

/* ====================================================================== */
/* Key switching                                                            */
/* ====================================================================== */

/* Context::MultKeyNoModDown (Context.cpp:609-712): ModUp(x), key inner
 * product over the groups (PAccum), then accum[q] += P * add (CAccum). The
 * key is full size; limb t of the extended basis reads key limb
 * GetPrimeVector index = global prime index (Container view with
 * np_front_ignore = max_ter - num_ter and the extra_ gap). */
int or_mult_key_no_moddown(const or_param *p, int level, const uint32_t *x,
                           const uint32_t *add_b, const uint32_t *add_a,
                           const uint32_t *evk_b, const uint32_t *evk_a,
                           uint32_t *accum_b, uint32_t *accum_a) {
  const int n = p->n, alpha = p->alpha;
  if (level < 0 || level > p->max_level) return -1;
  int nm = p->lv_main[level], nt = p->lv_ter[level];
  int num_q = nm + nt, tot = num_q + alpha;
  int pad = p->max_ter - nt, padded = num_q + pad;
  int beta = or_modup_beta(p, level);
  if (beta > p->dnum) return -1;
  size_t key_stride = (size_t)(p->L + alpha) * n;
  uint32_t *mu = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)beta * tot * n);
  or_modup(p, level, x, mu);
  int *lp = (int *)malloc(sizeof(int) * tot);
  level_prime_list(p, nm, nt, alpha, lp);
#pragma omp parallel for schedule(static) num_threads(g_num_threads)
  for (int t = 0; t < tot; t++) {
    int pi = lp[t];
    uint32_t q = p->q[pi];
    int32_t qinv = p->qinv[pi];
    for (int k = 0; k < n; k++) {
      uint32_t rb = 0, ra = 0;
      for (int g = 0; g < beta; g++) {
        int e = (g + 1) * alpha < padded ? (g + 1) * alpha : padded;
        if (e <= pad) continue;
        uint32_t m = mu[((size_t)g * tot + t) * n + k];
        rb = add_mod(rb, or_mult_mont(m, evk_b[g * key_stride +
                                               (size_t)pi * n + k],
                                      q, qinv),
                     q);
        ra = add_mod(ra, or_mult_mont(m, evk_a[g * key_stride +
                                               (size_t)pi * n + k],
                                      q, qinv),
                     q);
      }
      if (t < num_q) {
        uint32_t pp = p->p_prod_mont[pi];
        if (add_b)
          rb = add_mod(rb, or_mult_mont(pp, add_b[(size_t)t * n + k], q, qinv),
                       q);
        if (add_a)
          ra = add_mod(ra, or_mult_mont(pp, add_a[(size_t)t * n + k], q, qinv),
                       q);
      }
      accum_b[(size_t)t * n + k] = rb;
      accum_a[(size_t)t * n + k] = ra;
    }
  }
  free(mu);
  free(lp);
  return 0;
}

/* Context::MultKey at level -1: sparse-secret encapsulation's dense ->
 * sparse key switch (Context.cpp:567-607; AdjustLevelForMultKey turns level
 * 0 into -1 when the key's aux count differs from alpha). The level -1
 * ModSwitchHandler (ModSwitch.cu:175-232) has beta 1, no front padding and
 * num_aux = num_q auxiliary primes, the first ones (Parameter::
 * LevelToNP(-1, n) and GetPrimeVector, Parameter.cu:287-296, :318-325):
 *   ModUp(a)  (ModSwitch.cu:429-509): a's q limbs as they are, the aux limbs
 *             by BaseConv of INTT(a) x c1 and a plain forward NTT;
 *   PAccum    (Context.cpp:609-646, prime offset 0): acc = ModUp(a) (*) key
 *             over the key's own NP (q limbs, then aux limbs);
 *   CAccum    (Context.cpp:700-711): acc_b[q] += p_prod_dts (*) b, with
 *             p_prod_dts = prod aux mod q in Montgomery form (:109-123);
 *   ModDown   (ModSwitch.cu:566-684) of acc_b and acc_a: (acc[q] -
 *             NTT(BaseConv(INTT(acc[aux]) x c1))) (*) (prod aux)^-1 R.
 * key_b / key_a: 2 num_q limbs each. Returns -1 for an invalid base. */
int or_mult_key_short_base(const or_param *p, int num_main, int num_ter,
                           const uint32_t *b, const uint32_t *a,
                           const uint32_t *key_b, const uint32_t *key_a,
                           uint32_t *out_b, uint32_t *out_a) {
  const int n = p->n, nq = num_main + num_ter, tot = 2 * nq;
  if (num_main < 0 || num_ter < 0 || nq < 1 || nq > p->alpha ||
      num_main > p->max_main || num_ter > p->max_ter)
    return -1;
  const size_t qs = (size_t)nq * n;
  int *lp = (int *)malloc(sizeof(int) * tot);
  level_prime_list(p, num_main, num_ter, nq, lp);
  const int *aux = lp + nq;
  uint32_t *c1 = (uint32_t *)malloc(sizeof(uint32_t) * nq);
  int32_t *table = (int32_t *)malloc(sizeof(int32_t) * (size_t)nq * nq);
  int32_t *v = (int32_t *)malloc(sizeof(int32_t) * qs);
  uint32_t *mu = (uint32_t *)malloc(sizeof(uint32_t) * 2 * qs);
  uint32_t *acc = (uint32_t *)malloc(sizeof(uint32_t) * 4 * qs);
  uint32_t *y = (uint32_t *)malloc(sizeof(uint32_t) * qs);
  uint32_t *pp = (uint32_t *)malloc(sizeof(uint32_t) * nq);
  uint32_t *inv_prod = (uint32_t *)malloc(sizeof(uint32_t) * nq);

  /* ModUp: the one group (every q limb) to the aux primes */
  or_bconv_tables(p, lp, nq, aux, nq, NULL, 0, c1, table);
  or_intt_mult_const(p, lp, nq, a, c1, 1, v);
  memcpy(mu, a, sizeof(uint32_t) * qs);
  or_bconv(p, lp, nq, v, aux, nq, table, mu + qs);
  for (int j = 0; j < nq; j++) or_ntt_limb(p, aux[j], mu + qs + (size_t)j * n, 0);

  /* key inner product and the P' b fold */
  for (int i = 0; i < nq; i++) {
    uint32_t q = p->q[lp[i]], prod = 1;
    for (int j = 0; j < nq; j++) prod = or_mulmod(prod, p->q[aux[j]] % q, q);
    pp[i] = or_to_mont(prod, q);
    inv_prod[i] = or_to_mont(or_invmod(prod, q), q);
  }
  uint32_t *acc_b = acc, *acc_a = acc + 2 * qs;
  for (int t = 0; t < tot; t++) {
    const uint32_t q = p->q[lp[t]];
    const int32_t qinv = p->qinv[lp[t]];
    for (int k = 0; k < n; k++) {
      const size_t o = (size_t)t * n + k;
      uint32_t rb = or_mult_mont(mu[o], key_b[o], q, qinv);
      const uint32_t ra = or_mult_mont(mu[o], key_a[o], q, qinv);
      if (t < nq) rb = add_mod(rb, or_mult_mont(pp[t], b[o], q, qinv), q);
      acc_b[o] = rb;
      acc_a[o] = ra;
    }
  }

  /* ModDown of both accumulators: aux -> q */
  or_bconv_tables(p, aux, nq, lp, nq, NULL, 0, c1, table);
  uint32_t *outs[2] = {out_b, out_a};
  const uint32_t *accs[2] = {acc_b, acc_a};
  for (int w = 0; w < 2; w++) {
    or_intt_mult_const(p, aux, nq, accs[w] + qs, c1, 1, v);
    or_bconv(p, aux, nq, v, lp, nq, table, y);
    for (int i = 0; i < nq; i++) {
      const uint32_t q = p->q[lp[i]];
      const int32_t qinv = p->qinv[lp[i]];
      uint32_t *yl = y + (size_t)i * n;
      or_ntt_limb(p, lp[i], yl, 0);
      for (int k = 0; k < n; k++) {
        const size_t o = (size_t)i * n + k;
        outs[w][o] = or_mult_mont(sub_mod(accs[w][o], yl[k], q), inv_prod[i], q, qinv);
      }
    }
  }
  free(lp);
  free(c1);
  free(table);
  free(v);
  free(mu);
  free(acc);
  free(y);
  free(pp);
  free(inv_prod);
  return 0;
}

/* BootContext::ModUpToMax (BootContext.cpp:186-268): the level -1
 * ciphertext (short base (nm, nt), 1 or 2 primes) raised to every q prime
 * of the top level.
 *   sse: MultKey at level -1 with the dense -> sparse key first;
 *   per polynomial: INTTAndMultConst with ((prod_{j!=i} q_j) N)^-1 (plain;
 *     BootContext.cpp:93-107), ModUpToMax1/2 (ElementWise.cu:579-623: one
 *     prime s0 mod t, two primes s0 q1 + s1 q0 brought into (-q0 q1 / 2,
 *     q0 q1 / 2]), NTT with Montgomery conversion; with sse a's lift also
 *     covers the alpha auxiliary primes;
 *   sse: PMult by the sparse -> dense key's block 0, acc_b[q] += P (*) b
 *     (CAccum with GetPProd), ModDown of both (the top level's handler).
 * Keys: dts_* 2 (nm + nt) limbs (short base, then auxiliary); std_* the top
 * level's q limbs then alpha auxiliary limbs. Output: the top level's q
 * limbs. Returns -1 for an unsupported base. */
int or_mod_up_to_max(const or_param *p, int nm, int nt, int sse,
                     const uint32_t *b, const uint32_t *a,
                     const uint32_t *dts_b, const uint32_t *dts_a,
                     const uint32_t *std_b, const uint32_t *std_a,
                     uint32_t *out_b, uint32_t *out_a) {
  const int n = p->n, k = nm + nt, alpha = p->alpha, top = p->max_level;
  if (k < 1 || k > 2) return -1;
  const int tm = p->lv_main[top], tt = p->lv_ter[top], tq = tm + tt;
  const size_t ks = (size_t)k * n, tqs = (size_t)tq * n;
  uint32_t *w = (uint32_t *)malloc(sizeof(uint32_t) * 2 * ks);
  if (sse) {
    if (or_mult_key_short_base(p, nm, nt, b, a, dts_b, dts_a, w, w + ks)) {
      free(w);
      return -1;
    }
  } else {
    memcpy(w, b, sizeof(uint32_t) * ks);
    memcpy(w + ks, a, sizeof(uint32_t) * ks);
  }
  int base[2], *lp = (int *)malloc(sizeof(int) * (tq + alpha));
  uint32_t c[2];
  for (int i = 0; i < k; i++) base[i] = or_np_prime(p, nm, nt, i);
  for (int i = 0; i < k; i++) {
    const uint32_t q = p->q[base[i]];
    uint32_t v = 1;
    for (int j = 0; j < k; j++)
      if (j != i) v = or_mulmod(v, p->q[base[j]] % q, q);
    c[i] = or_invmod(or_mulmod(v, (uint32_t)n % q, q), q);
  }
  level_prime_list(p, tm, tt, alpha, lp);
  int32_t *coef = (int32_t *)malloc(sizeof(int32_t) * ks);
  uint32_t *lift = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (tqs + (size_t)alpha * n));
  uint32_t *lb = lift, *la = lift + tqs + (size_t)alpha * n;
  const int64_t q0 = p->q[base[0]], q1 = k == 2 ? p->q[base[1]] : 1, qq = q0 * q1;
  for (int poly = 0; poly < 2; poly++) {
    const int nl = tq + ((sse && poly == 1) ? alpha : 0);
    uint32_t *dst = poly == 0 ? lb : la;
    or_intt_mult_const(p, base, k, w + poly * ks, c, 0, coef);
    for (int t = 0; t < nl; t++) {
      const int64_t m = p->q[lp[t]];
      for (int x = 0; x < n; x++) {
        int64_t v;
        if (k == 2) {
          v = (int64_t)(uint32_t)coef[x] * q1 + (int64_t)(uint32_t)coef[n + x] * q0;
          if (v >= qq) v -= qq;
          if (v > qq / 2) v -= qq;
        } else {
          v = (uint32_t)coef[x];
        }
        int64_t r = v % m;
        if (r < 0) r += m;
        dst[(size_t)t * n + x] = (uint32_t)r;
      }
      or_ntt_limb(p, lp[t], dst + (size_t)t * n, 1);
    }
  }
  if (!sse) {
    memcpy(out_b, lb, sizeof(uint32_t) * tqs);
    memcpy(out_a, la, sizeof(uint32_t) * tqs);
  } else {
    const int tot = tq + alpha;
    uint32_t *acc = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)tot * n);
    uint32_t *acc_b = acc, *acc_a = acc + (size_t)tot * n;
    for (int t = 0; t < tot; t++) {
      const uint32_t q = p->q[lp[t]];
      const int32_t qinv = p->qinv[lp[t]];
      for (int x = 0; x < n; x++) {
        const size_t o = (size_t)t * n + x;
        uint32_t rb = or_mult_mont(std_b[o], la[o], q, qinv);
        acc_a[o] = or_mult_mont(std_a[o], la[o], q, qinv);
        if (t < tq) rb = add_mod(rb, or_mult_mont(lb[o], p->p_prod_mont[lp[t]], q, qinv), q);
        acc_b[o] = rb;
      }
    }
    int rc = or_moddown(p, top, OR_MODDOWN, acc_b, out_b);
    if (!rc) rc = or_moddown(p, top, OR_MODDOWN, acc_a, out_a);
    free(acc);
    if (rc) {
      free(w);
      free(lp);
      free(coef);
      free(lift);
      return -1;
    }
  }
  free(w);
  free(lp);
  free(coef);
  free(lift);
  return 0;
}

/* Context::HMult (Context.cpp:862-871) = Mult (Tensor, :316-330) then
 * RelinearizeRescale (:714-737) or Relinearize -> MultKey (:584-607). */
int or_hmult(const or_param *p, int level, const uint32_t *b1,
             const uint32_t *a1, const uint32_t *b2, const uint32_t *a2,
             const uint32_t *evk_b, const uint32_t *evk_a, int rescale,
             uint32_t *out_b, uint32_t *out_a) {
  const int n = p->n, alpha = p->alpha;
  if (level < 0 || level > p->max_level || (rescale && level == 0)) return -1;
  int nm = p->lv_main[level], nt = p->lv_ter[level];
  int num_q = nm + nt, tot = num_q + alpha;
  size_t qs = (size_t)num_q * n;
  int *lp = (int *)malloc(sizeof(int) * num_q);
  level_prime_list(p, nm, nt, 0, lp);
  uint32_t *d = (uint32_t *)malloc(sizeof(uint32_t) * 3 * qs);
  or_tensor(p, lp, num_q, b1, a1, b2, a2, d, d + qs, d + 2 * qs);
  uint32_t *acc = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)tot * n);
  int rc = or_mult_key_no_moddown(p, level, d + 2 * qs, d, d + qs, evk_b,
                                  evk_a, acc, acc + (size_t)tot * n);
  if (rc == 0) {
    int kind = rescale ? OR_MODDOWN_RESCALE : OR_MODDOWN;
    rc = or_moddown(p, level, kind, acc, out_b);
    if (rc == 0) rc = or_moddown(p, level, kind, acc + (size_t)tot * n, out_a);
  }
  free(lp);
  free(d);
  free(acc);
  return rc;
}

/* MultKey on ax with accum.bx += P*bx (Context.cpp:685-699), ModDown both,
 * then Permute by the Galois factor g (Context.cpp:502-527). */
static int hrot_galois(const or_param *p, int level, const uint32_t *b,
                       const uint32_t *a, const uint32_t *rk_b,
                       const uint32_t *rk_a, uint32_t g, uint32_t *out_b,
                       uint32_t *out_a) {
  const int n = p->n, alpha = p->alpha;
  int nm = p->lv_main[level], nt = p->lv_ter[level];
  int num_q = nm + nt, tot = num_q + alpha;
  size_t qs = (size_t)num_q * n;
  uint32_t *acc = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)tot * n);
  uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * 2 * qs);
  int rc = or_mult_key_no_moddown(p, level, a, b, NULL, rk_b, rk_a, acc,
                                  acc + (size_t)tot * n);
  if (rc == 0) rc = or_moddown(p, level, OR_MODDOWN, acc, tmp);
  if (rc == 0)
    rc = or_moddown(p, level, OR_MODDOWN, acc + (size_t)tot * n, tmp + qs);
  if (rc == 0) {
    or_permute(p, num_q, g, tmp, out_b);
    or_permute(p, num_q, g, tmp + qs, out_a);
  }
  free(acc);
  free(tmp);
  return rc;
}

/* Context::HRot (Context.cpp:776-791): rotation by rot (reduced mod N/2;
 * 0 copies), Galois factor 5^rot. */
int or_hrot(const or_param *p, int level, const uint32_t *b,
            const uint32_t *a, const uint32_t *rk_b, const uint32_t *rk_a,
            int rot, uint32_t *out_b, uint32_t *out_a) {
  if (level < 0 || level > p->max_level) return -1;
  const int slots = p->n / 2;
  size_t qs = (size_t)(p->lv_main[level] + p->lv_ter[level]) * p->n;
  rot %= slots;
  if (rot < 0) rot += slots;
  if (rot == 0) {
    memcpy(out_b, b, sizeof(uint32_t) * qs);
    memcpy(out_a, a, sizeof(uint32_t) * qs);
    return 0;
  }
  return hrot_galois(p, level, b, a, rk_b, rk_a, or_galois_factor(p, rot),
                     out_b, out_a);
}

/* Context::HConj (Context.cpp:793-798): MultKey then Permute by 2N - 1 */
int or_hconj(const or_param *p, int level, const uint32_t *b,
             const uint32_t *a, const uint32_t *ck_b, const uint32_t *ck_a,
             uint32_t *out_b, uint32_t *out_a) {
  if (level < 0 || level > p->max_level) return -1;
  return hrot_galois(p, level, b, a, ck_b, ck_a, or_galois_factor(p, -1),
                     out_b, out_a);
}

/* HoistHandler baby step for one rotation: MultKeyNoModDown(a) over the
 * key groups plus P * b on the q limbs (PseudoModUp, Hoist.cu:627-651), then
 * BSFusedKernel's scatter (Hoist.cu:41-75): coefficient x of every limb goes
 * to brv17((brv17(x) + 1) * g - 1) with g = 5^(N/2 - rot) (Hoist.cu:300-302).
 * The result stays in the Q*P basis. */
int or_hoisted_baby_step(const or_param *p, int level, const uint32_t *b,
                         const uint32_t *a, const uint32_t *rk_b,
                         const uint32_t *rk_a, int rot, uint32_t *out_b,
                         uint32_t *out_a) {
  const int n = p->n, alpha = p->alpha, lb = p->log_n + 1;
  const uint32_t mask = (2u << p->log_n) - 1;
  if (level < 0 || level > p->max_level || rot < 1 || rot >= n / 2) return -1;
  int nm = p->lv_main[level], nt = p->lv_ter[level];
  int tot = nm + nt + alpha;
  uint32_t *acc = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)tot * n);
  int rc = or_mult_key_no_moddown(p, level, a, b, NULL, rk_b, rk_a, acc,
                                  acc + (size_t)tot * n);
  if (rc == 0) {
    const uint32_t g = or_galois_factor(p, n / 2 - rot);
#pragma omp parallel for schedule(static) num_threads(g_num_threads)
    for (int t = 0; t < tot; t++) {
      for (int x = 0; x < n; x++) {
        uint32_t e = or_bitrev((uint32_t)x, lb) + 1;
        uint32_t d = or_bitrev(((e * g) - 1) & mask, lb);
        out_b[(size_t)t * n + d] = acc[(size_t)t * n + x];
        out_a[(size_t)t * n + d] = acc[(size_t)(tot + t) * n + x];
      }
    }
  }
  free(acc);
  return rc;
}
