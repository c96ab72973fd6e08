/*
 * CPU oracle for the CKKS RNS/NTT key-switching path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it. The product library never links or calls it.
 *
 * It is a scalar C restatement of the reference's (scale-snu/cheddar-fhe)
 * algorithms for 32-bit words; each function cites the reference file:line
 * it follows. All arithmetic is exact integer arithmetic on u32 residues of
 * primes q < 2^31.
 *
 * Layout conventions (same as the reference):
 *   - polynomials are limb-major: data[limb * N + coeff], u32 in [0, q)
 *   - limb order of an "NP" (num_main, num_ter, num_aux) is
 *     [terminal primes reversed | main primes | auxiliary primes]
 *     (src/core/Parameter.cu:103-111, Parameter.cu:327-340)
 *   - every NTT-domain polynomial holds x * 2^32 mod q (Montgomery form)
 *
 * Parity status of each function group is recorded at the top of
 * ckks_oracle.c and pinned by tests/test_oracle_*.py.
 */
#ifndef CKKS_ORACLE_H_
#define CKKS_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Scalar number theory (include/common/PrimeUtils.h, Basic.cuh)           */
/* ---------------------------------------------------------------------- */
uint32_t or_mulmod(uint32_t a, uint32_t b, uint32_t q);
uint32_t or_powmod(uint32_t a, int64_t e, uint32_t q);
uint32_t or_invmod(uint32_t a, uint32_t q);
uint32_t or_find_generator(uint32_t q);
uint32_t or_primitive_root(int64_t m, uint32_t q);
int32_t or_inv_mod_base(uint32_t q);
uint32_t or_to_mont(uint32_t a, uint32_t q);
int32_t or_mont_reduce_lazy(int64_t a, uint32_t q, int32_t qinv);
uint32_t or_mont_reduce(int64_t a, uint32_t q, int32_t qinv);
uint32_t or_mult_mont(uint32_t a, uint32_t b, uint32_t q, int32_t qinv);
int32_t or_normalize(uint32_t a, uint32_t q);
uint32_t or_bitrev(uint32_t i, int bits);

/* ---------------------------------------------------------------------- */
/* Parameters (src/core/Parameter.cu:20-371)                               */
/* ---------------------------------------------------------------------- */
typedef struct or_param or_param;

/* level_cfg holds n_levels pairs (num_main, num_ter). Returns NULL on an
 * invalid parameter set (the reference exits via AssertTrue instead). */
or_param *or_param_create(int log_n, const uint32_t *main_primes, int n_main,
                          const uint32_t *ter_primes, int n_ter,
                          const uint32_t *aux_primes, int n_aux,
                          const int *level_cfg, int n_levels);
void or_param_destroy(or_param *p);
/* out[0..9] = log_n, n, max_main, max_ter, alpha, L, dnum, max_level,
 *             num_primes (L + alpha), 0 */
void or_param_info(const or_param *p, int *out);
/* all primes in global order [ter reversed | main | aux] */
uint32_t or_prime(const or_param *p, int pidx);
int32_t or_prime_qinv(const or_param *p, int pidx);
uint32_t or_prime_psi(const or_param *p, int pidx);
/* level -> (num_main, num_ter); returns 0 on success */
int or_level_np(const or_param *p, int level, int *num_main, int *num_ter);
/* (num_main, num_ter) -> level, or -1 when absent */
int or_np_level(const or_param *p, int num_main, int num_ter);
/* global prime index of limb `limb` of NP (num_main, num_ter, num_aux) */
int or_np_prime(const or_param *p, int num_main, int num_ter, int limb);
/* Galois factor 5^r mod 2N (Parameter.cu:93-101); r == -1 -> 2N - 1 */
uint32_t or_galois_factor(const or_param *p, int r);
/* per-prime tables: psi_rev and psi_inv_rev in Montgomery form (N entries) */
const uint32_t *or_psi_rev_mont(const or_param *p, int pidx);
const uint32_t *or_psi_inv_rev_mont(const or_param *p, int pidx);
/* P mod q_i in Montgomery form, indexed by global q prime index */
uint32_t or_p_prod_mont(const or_param *p, int pidx);

/* number of worker threads (OpenMP) used by the polynomial routines */
void or_set_num_threads(int n);
int or_get_num_threads(void);

/* ---------------------------------------------------------------------- */
/* Seeded portable RNG: SplitMix64 seeding -> xoshiro256**                 */
/* ---------------------------------------------------------------------- */
uint64_t or_splitmix64(uint64_t *state);
/* uniform residues in [0, q) for each limb of NP, rejection sampled */
void or_fill_uniform(const or_param *p, int num_main, int num_ter,
                     int num_aux, uint64_t seed, uint32_t *out);
/* uniform residues for an explicit list of global prime indices */
void or_fill_uniform_primes(const or_param *p, const int *pidx, int nlimbs,
                            uint64_t seed, uint32_t *out);

/* ---------------------------------------------------------------------- */
/* NTT (src/core/NTT.cu, include/core/NTTUtils.cuh)                        */
/* ---------------------------------------------------------------------- */
/* In-place forward negacyclic NTT of one limb, canonical [0,q) output.
 * mont_conv multiplies the input by 2^32 first (NTT(..., true)). */
void or_ntt_limb(const or_param *p, int pidx, uint32_t *a, int mont_conv);
/* In-place inverse NTT of one limb WITHOUT the N^-1 factor, [0,q) output */
void or_intt_raw_limb(const or_param *p, int pidx, uint32_t *a);
/* NTTHandler::NTT over every limb of an NP */
void or_ntt(const or_param *p, int num_main, int num_ter, int num_aux,
            uint32_t *data, int mont_conv);
/* NTTHandler::INTT over every limb of an NP. mont_conv = 1 leaves the
 * Montgomery form (x * N^-1 * R^-1), 0 keeps it (NTT.cu:673-741). */
void or_intt(const or_param *p, int num_main, int num_ter, int num_aux,
             uint32_t *data, int mont_conv);
/* INTTAndMultConst (NTT.cu:743-813): out = INTT_raw(src) (x) c mod q in
 * Montgomery product form; normalize selects the centred output range. */
void or_intt_mult_const(const or_param *p, const int *pidx, int nlimbs,
                        const uint32_t *src, const uint32_t *consts,
                        int normalize, int32_t *out);

/* ---------------------------------------------------------------------- */
/* Fast base conversion (src/core/ModSwitch.cu:23-171, :320-366)           */
/* ---------------------------------------------------------------------- */
/* Builds the two constant tables exactly as PopulateModSwitchConstants:
 * const1[i] (s entries) and table[m * s + i] (d * s entries, signed).
 * restore primes (global indices) multiply const1 (BitPacker restore). */
void or_bconv_tables(const or_param *p, const int *src_pidx, int s,
                     const int *dst_pidx, int d, const int *restore_pidx,
                     int n_restore, uint32_t *const1, int32_t *table);
/* y_m = ReduceMontgomery(sum_i src[i] * table[m][i]) for each coeff */
void or_bconv(const or_param *p, const int *src_pidx, int s,
              const int32_t *src, const int *dst_pidx, int d,
              const int32_t *table, uint32_t *dst);

/* ---------------------------------------------------------------------- */
/* ModUp / ModDown (src/core/ModSwitch.cu:175-318, :429-684)               */
/* ---------------------------------------------------------------------- */
/* number of ModUp groups (beta) at a level (ModSwitch.cu:181-185) */
int or_modup_beta(const or_param *p, int level);
/* ModUp of a (num_q)-limb NTT-domain poly. out holds beta blocks of
 * (num_q + alpha) limbs; blocks of groups that lie fully in the padding are
 * zero-filled. */
void or_modup(const or_param *p, int level, const uint32_t *src,
              uint32_t *out);
enum { OR_MODDOWN = 0, OR_RESCALE = 1, OR_MODDOWN_RESCALE = 2 };
/* ModDownWorker (ModSwitch.cu:566-684) */
int or_moddown(const or_param *p, int level, int kind, const uint32_t *src,
               uint32_t *dst);

/* ---------------------------------------------------------------------- */
/* Element-wise (src/core/ElementWise.cu)                                  */
/* ---------------------------------------------------------------------- */
void or_add(const or_param *p, const int *pidx, int nlimbs,
            const uint32_t *x, const uint32_t *y, uint32_t *out);
void or_sub(const or_param *p, const int *pidx, int nlimbs,
            const uint32_t *x, const uint32_t *y, uint32_t *out);
void or_neg(const or_param *p, const int *pidx, int nlimbs,
            const uint32_t *x, uint32_t *out);
void or_mult(const or_param *p, const int *pidx, int nlimbs,
             const uint32_t *x, const uint32_t *y, uint32_t *out);
void or_mult_const(const or_param *p, const int *pidx, int nlimbs,
                   const uint32_t *x, const uint32_t *consts, uint32_t *out);
void or_tensor(const or_param *p, const int *pidx, int nlimbs,
               const uint32_t *b1, const uint32_t *a1, const uint32_t *b2,
               const uint32_t *a2, uint32_t *d0, uint32_t *d1, uint32_t *d2);
void or_permute(const or_param *p, int nlimbs, uint32_t galois_factor,
                const uint32_t *src, uint32_t *dst);

/* ---------------------------------------------------------------------- */
/* Key switching (src/core/Context.cpp:584-871)                            */
/* ---------------------------------------------------------------------- */
/* Evaluation keys are stored at full size, as EvaluationKey<word>: evk_b and
 * evk_a each hold dnum blocks of (L + alpha) limbs in global prime order. */
/* MultKeyNoModDown: accum_{b,a} over (num_q + alpha) limbs.
 * x is the polynomial to key-switch (rx for relinearization, ax for
 * rotations); add_b/add_a (num_q limbs, may be NULL) are folded in as
 * accum[q] += P * add. */
int or_mult_key_no_moddown(const or_param *p, int level, const uint32_t *x,
                           const uint32_t *add_b, const uint32_t *add_a,
                           const uint32_t *evk_b, const uint32_t *evk_a,
                           uint32_t *accum_b, uint32_t *accum_a);
/* MultKey at level -1 (the dense -> sparse switch of sparse-secret
 * encapsulation) on the short base (num_main, num_ter) with as many
 * auxiliary primes; key_b / key_a hold 2 num_q limbs (short base, then
 * auxiliary) */
int or_mult_key_short_base(const or_param *p, int num_main, int num_ter,
                           const uint32_t *b, const uint32_t *a,
                           const uint32_t *key_b, const uint32_t *key_a,
                           uint32_t *out_b, uint32_t *out_a);
/* BootContext::ModUpToMax from the short base (nm, nt) (1 or 2 primes) to
 * the top level's q limbs; with sse between the dense -> sparse switch
 * (dts_*: 2 (nm + nt) limbs) and the sparse -> dense switch at the top
 * level (std_*: the top level's q limbs then alpha auxiliary limbs) */
int or_mod_up_to_max(const or_param *p, int nm, int nt, int sse,
                     const uint32_t *b, const uint32_t *a,
                     const uint32_t *dts_b, const uint32_t *dts_a,
                     const uint32_t *std_b, const uint32_t *std_a,
                     uint32_t *out_b, uint32_t *out_a);
/* HMult = Tensor + RelinearizeRescale (rescale=1) or Relinearize (0).
 * Output limbs: num_q(level - 1) when rescaling, num_q(level) otherwise. */
int or_hmult(const or_param *p, int level, const uint32_t *b1,
             const uint32_t *a1, const uint32_t *b2, const uint32_t *a2,
             const uint32_t *evk_b, const uint32_t *evk_a, int rescale,
             uint32_t *out_b, uint32_t *out_a);
/* HRot = MultKey + Permute (Context.cpp:776-791) */
int or_hrot(const or_param *p, int level, const uint32_t *b,
            const uint32_t *a, const uint32_t *rk_b, const uint32_t *rk_a,
            int rot, uint32_t *out_b, uint32_t *out_a);

/* HConj = MultKey + Permute(2N - 1) (Context.cpp:793-798) */
int or_hconj(const or_param *p, int level, const uint32_t *b,
             const uint32_t *a, const uint32_t *ck_b, const uint32_t *ck_a,
             uint32_t *out_b, uint32_t *out_a);
/* Hoisted baby step for one rotation (HoistHandler::EvaluateBabyStep,
 * src/extension/Hoist.cu:578-691, BSFusedKernel :21-78): out over
 * (num_q + alpha) limbs, rot in [1, N/2). */
int or_hoisted_baby_step(const or_param *p, int level, const uint32_t *b,
                         const uint32_t *a, const uint32_t *rk_b,
                         const uint32_t *rk_a, int rot, uint32_t *out_b,
                         uint32_t *out_a);

#ifdef __cplusplus
}
#endif
#endif /* CKKS_ORACLE_H_ */
