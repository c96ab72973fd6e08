/*
 * cheddar_amd.h -- C ABI of the MI355X (gfx950) CKKS key-switching core.
 *
 * Drop-in boundary for the handler layer of scale-snu/cheddar-fhe (32-bit
 * word path). Each entry point names the reference interface it replaces.
 *
 * Conventions (identical to the reference):
 *   - a polynomial is limb-major u32 [limb][N], N = 2^16, residues in [0,q);
 *   - limb order of an NP (num_main, num_ter, num_aux) is
 *     [terminal primes reversed | main primes | auxiliary primes]
 *     (src/core/Parameter.cu:103-111);
 *   - NTT-domain data hold x * 2^32 mod q (Montgomery form); products of
 *     stored values are Montgomery products x * y * 2^-32;
 *   - evaluation keys are full size (EvaluationKey, include/core/Container.h:
 *     288-345): dnum blocks, each a (L + alpha)-limb polynomial in global
 *     prime order, passed as host arrays of dnum device pointers;
 *   - `src_gap` is the reference's DvView extra_ (include/common/PtrList.h:
 *     37-63) in limbs: aux limb a of a source sits at limb num_q + a + gap.
 *
 * Every pointer to polynomial data is a device pointer; `stream` is a
 * hipStream_t (NULL = default stream). Entry points enqueue work and return
 * without synchronising. Operations of one context share one workspace:
 * calls on one context must not overlap in time (serialise them on one
 * stream). Errors return a non-zero status and set a message readable with
 * ck_last_error(); nothing exits the process (the reference's AssertTrue
 * calls std::exit, include/common/Assert.h:9-21).
 */
#ifndef CHEDDAR_AMD_H_
#define CHEDDAR_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CK_OK 0
#define CK_ERR_INVALID_ARG 1
#define CK_ERR_UNSUPPORTED 2
#define CK_ERR_HIP 3
#define CK_ERR_OUT_OF_MEMORY 4

/* ModDown flavours (ModSwitchHandler::ModDown / Rescale / ModDownAndRescale,
 * include/core/ModSwitch.h:38-44) */
#define CK_MODDOWN 0
#define CK_RESCALE 1
#define CK_MODDOWN_RESCALE 2

typedef struct ck_context ck_context;

/* Message of the last failing call on this thread. */
const char *ck_last_error(void);
/* ABI version (major * 100 + minor). */
int ck_version(void);

/* ---- Parameter + Context -------------------------------------------- */

/* Replaces Parameter<word>::Parameter (include/core/Parameter.h:42-47,
 * src/core/Parameter.cu:20-193) together with Context<word>::Create
 * (include/core/Context.h:70, src/core/Context.cpp:62-124): validates the
 * primes and the level table, builds twiddles, ModUp/ModDown constants and
 * the P table, and allocates the workspace. level_config holds n_levels
 * pairs (num_main, num_ter), as the parameter JSON's "level_config". */
int ck_context_create(int log_degree, const uint32_t *main_primes, int n_main,
                      const uint32_t *ter_primes, int n_ter,
                      const uint32_t *aux_primes, int n_aux,
                      const int *level_config, int n_levels,
                      ck_context **out);
int ck_context_destroy(ck_context *ctx);
/* out[0..9] = log_n, N, max_main, max_ter, alpha, L, dnum, max_level,
 * num_primes, workspace bytes / 2^20 */
int ck_context_info(const ck_context *ctx, int *out);
/* Level -1 for sparse-secret encapsulation: the short base (Parameter's
 * short_base_np_, src/core/Parameter.cu:139-147: level 0's primes less the
 * additional base) with as many auxiliary primes (GetSSENumAux,
 * :271-274), the level -1 ModSwitchHandler (src/core/ModSwitch.cu:175-232)
 * and P' mod q (p_prod_dts_, src/core/Context.cpp:109-123). Afterwards
 * ck_level_to_np, ck_modup_beta, ck_modup, ck_moddown (kind 0),
 * ck_mult_key_no_moddown and ck_mult_key accept level -1 (Context::MultKey
 * with the DtS handler, src/core/Context.cpp:567-607); its key blocks hold
 * the short base's limbs followed by its auxiliary limbs (2 num_q limbs).
 * num_main + num_ter in 1..alpha; a second call with the same base is a
 * no-op. */
int ck_context_set_short_base(ck_context *ctx, int num_main, int num_ter);
/* Parameter::LevelToNP (src/core/Parameter.cu:287-296) */
int ck_level_to_np(const ck_context *ctx, int level, int *num_main,
                   int *num_ter);
/* Parameter::NPToLevel (src/core/Parameter.cu:298-315); -1 when absent */
int ck_np_to_level(const ck_context *ctx, int num_main, int num_ter);
/* beta of ModSwitchHandler at a level (src/core/ModSwitch.cu:181-185) */
int ck_modup_beta(const ck_context *ctx, int level);
/* Parameter::GetGaloisFactor (src/core/Parameter.cu:195-200) and
 * ElementWiseHandler::PermuteAmountToGaloisFactor (ElementWise.cu:955-964):
 * rot in [0, N/2] -> 5^rot mod 2N, rot == -1 -> 2N - 1 (conjugation) */
int ck_galois_factor(const ck_context *ctx, int rot, uint32_t *out);
/* the prime of global index pidx ([ter reversed | main | aux]) */
int ck_prime(const ck_context *ctx, int pidx, uint32_t *out);

/* ---- NTTHandler (include/core/NTT.h:56-81) --------------------------- */

/* NTTHandler::NTT (src/core/NTT.cu:606-671): forward negacyclic NTT of every
 * limb of NP (num_main, num_ter, num_aux); montgomery_conversion multiplies
 * the input by 2^32 first. dst may equal src. */
int ck_ntt(ck_context *ctx, uint32_t *dst, const uint32_t *src, int num_main,
           int num_ter, int num_aux, int src_gap, int montgomery_conversion,
           void *stream);
/* NTTHandler::INTT (src/core/NTT.cu:673-741): inverse NTT times N^-1;
 * montgomery_conversion = 1 leaves the Montgomery form. dst may equal src. */
int ck_intt(ck_context *ctx, uint32_t *dst, const uint32_t *src, int num_main,
            int num_ter, int num_aux, int src_gap, int montgomery_conversion,
            void *stream);
/* NTTHandler::INTTAndMultConst (src/core/NTT.cu:743-813): dst = INTT(src)
 * (*) c per limb (Montgomery product, no N^-1), canonical or centred
 * (normalize: int32 in [-(q-1)/2, (q-1)/2]). consts is a HOST array of one
 * value per limb (the reference passes a device view; host constants keep
 * the call free of device-to-host reads). */
int ck_intt_mult_const(ck_context *ctx, uint32_t *dst, const uint32_t *src,
                       int num_main, int num_ter, int num_aux, int src_gap,
                       const uint32_t *consts, int normalize, void *stream);

/* ---- ModSwitchHandler (include/core/ModSwitch.h:38-44) ---------------- */

/* ModSwitchHandler::ModUp (src/core/ModSwitch.cu:429-509) at `level`: src
 * holds the level's num_q NTT-domain limbs; dst holds beta blocks of
 * (num_q + alpha) limbs, block g = group g's decomposition raised to Q*P.
 * Blocks of groups lying wholly in the terminal padding are not written. */
int ck_modup(ck_context *ctx, int level, uint32_t *dst, const uint32_t *src,
             void *stream);
/* ModSwitchHandler::ModDown / Rescale / ModDownAndRescale
 * (src/core/ModSwitch.cu:511-684) of npoly (1..4) polynomials at `level`.
 * kind = CK_MODDOWN: src (num_q + alpha) limbs -> dst num_q limbs;
 * CK_RESCALE: src num_q -> dst num_q(level - 1);
 * CK_MODDOWN_RESCALE: src (num_q + alpha) -> dst num_q(level - 1).
 * dst and src are host arrays of npoly device pointers; dst != src. */
int ck_moddown(ck_context *ctx, int level, int kind, int npoly,
               uint32_t *const *dst, const uint32_t *const *src,
               void *stream);

/* ---- ElementWiseHandler (include/core/ElementWise.h:52-116) ---------- */
/* npoly polynomials of NP (num_main, num_ter, num_aux) with no limb gap;
 * dst, x, y are host arrays of npoly (1..3) device pointers. */

/* ElementWiseHandler::Add (ElementWise.cu:688-716, kernel Sum :32-71) */
int ck_add(ck_context *ctx, int num_main, int num_ter, int num_aux, int npoly,
           uint32_t *const *dst, const uint32_t *const *x,
           const uint32_t *const *y, void *stream);
/* ElementWiseHandler::Sub (kernel Sub, ElementWise.cu:74-98) */
int ck_sub(ck_context *ctx, int num_main, int num_ter, int num_aux, int npoly,
           uint32_t *const *dst, const uint32_t *const *x,
           const uint32_t *const *y, void *stream);
/* ElementWiseHandler::Mult (kernel Mult, ElementWise.cu:120-143):
 * Montgomery product */
int ck_mult(ck_context *ctx, int num_main, int num_ter, int num_aux,
            int npoly, uint32_t *const *dst, const uint32_t *const *x,
            const uint32_t *const *y, void *stream);
/* ElementWiseHandler::Neg (kernel Neg, ElementWise.cu:100-118) */
int ck_neg(ck_context *ctx, int num_main, int num_ter, int num_aux, int npoly,
           uint32_t *const *dst, const uint32_t *const *x, void *stream);
/* ElementWiseHandler::MultConst (CPAccum<1,const>, ElementWise.cu:225-287):
 * x (*) c[limb]; consts is a DEVICE array of num_total values */
int ck_mult_const(ck_context *ctx, int num_main, int num_ter, int num_aux,
                  int npoly, uint32_t *const *dst, const uint32_t *const *x,
                  const uint32_t *consts, void *stream);
/* ElementWiseHandler::AddConst / SubConst / SubOppositeConst
 * (ElementWise.cu:145-223): mode 0: x + c[limb], 1: x - c[limb],
 * 2: c[limb] - x; consts is a DEVICE array of num_total values */
int ck_add_const(ck_context *ctx, int num_main, int num_ter, int num_aux,
                 int npoly, uint32_t *const *dst, const uint32_t *const *x,
                 const uint32_t *consts, int mode, void *stream);
/* ElementWiseHandler::MultImaginaryUnit (ElementWise.cu:625-658, :1673-1703):
 * x * X^(N/2) in the NTT domain (the slot-wise product by sqrt(-1)) */
int ck_mult_imaginary_unit(ck_context *ctx, int num_main, int num_ter,
                           int num_aux, int npoly, uint32_t *const *dst,
                           const uint32_t *const *x, void *stream);
/* ElementWiseHandler::ModUpToMax (ElementWise.cu:579-623, :1646-1671):
 * src = the coefficient-domain limbs of the base NP (base_main, base_ter),
 * one or two primes; dst = every main and terminal limb of the top level
 * followed by num_aux auxiliary limbs (coefficient domain, [0, q)). */
int ck_modup_to_max(ck_context *ctx, int base_main, int base_ter,
                    const uint32_t *src, uint32_t *dst, int num_aux,
                    void *stream);
/* The same lift between explicit global prime indices (the reference's
 * ModUpToMax kernels with free source / destination primes; used for the
 * two-prime key switches of sparse-secret encapsulation,
 * src/extension/BootContext.cpp:193-268): src = n_src (1 or 2)
 * coefficient-domain limbs s_i = [x (Q / q_i)^-1]_{q_i}; dst = x mod p for
 * each of n_dst (<= 128) destination primes, x centred for two primes. */
int ck_crt_lift(ck_context *ctx, const int *src_primes, int n_src,
                const int *dst_primes, int n_dst, const uint32_t *src, uint32_t *dst,
                void *stream);
/* ElementWiseHandler::PAccum (ElementWise.cu:1437-1643, CPAccum kernels
 * :225-366): for p < npoly (1 or 2),
 *   dst[p] = (accumulate ? dst[p] : 0) + sum_t x[t * npoly + p] (*) y[t]
 * with y[t] a polynomial; every operand holds the NP's limbs contiguously.
 * Any nterms >= 0 (nterms = 0 clears dst unless accumulating). */
int ck_paccum(ck_context *ctx, int num_main, int num_ter, int num_aux, int npoly,
              int nterms, uint32_t *const *dst, const uint32_t *const *x,
              const uint32_t *const *y, int accumulate, void *stream);
/* HoistHandler::GSFusedPAccum (src/extension/Hoist.cu:329-451, kernel
 * GSFusedKernel :80-118): the giant-step accumulations of a hoisted linear
 * transform in one pass, each baby step read once. For g < ngs, j in {0, 1}
 * (b, a):
 *   dst[2 g + j] = sum_{t < nbs, pt[g nbs + t] != NULL} x[2 t + j] (*) pt[g nbs + t]
 * over the limbs of the NP (every operand holds them contiguously);
 * 1 <= nbs <= 16, any ngs >= 0; no output may alias an input. */
int ck_gs_accum(ck_context *ctx, int num_main, int num_ter, int num_aux, int nbs,
                int ngs, uint32_t *const *dst, const uint32_t *const *x,
                const uint32_t *const *pt, void *stream);
/* ElementWiseHandler::PermuteAccum (ElementWise.cu:1002-1100, kernels
 * PermuteAccum / PermuteAccumAdd :394-501): a sum of NTT-domain
 * automorphisms, for j < npoly (1 or 2),
 *   dst[j] = (accumulate ? dst[j] : 0) + sum_{s < nsrc} sigma_{rots[s]}(src[s npoly + j])
 * with sigma_r as in ck_permute (r = 0: identity, r = -1: conjugation); a
 * NULL src entry adds nothing; no source may alias dst; every operand
 * holds the NP's limbs contiguously. */
int ck_permute_accum(ck_context *ctx, int num_main, int num_ter, int num_aux, int npoly,
                     uint32_t *const *dst, int nsrc, const uint32_t *const *src,
                     const int *rots, int accumulate, void *stream);
/* ElementWiseHandler::CAccum (same kernels, constant multipliers): as
 * ck_paccum with consts[t] a DEVICE array of one value per limb. */
int ck_caccum(ck_context *ctx, int num_main, int num_ter, int num_aux, int npoly,
              int nterms, uint32_t *const *dst, const uint32_t *const *x,
              const uint32_t *const *consts, int accumulate, void *stream);
/* ElementWiseHandler::Tensor (ElementWise.cu:921-953, kernel :504-546):
 * (d0, d1, d2) = (b1 b2, b1 a2 + a1 b2, a1 a2) on num_q limbs; with
 * b1 == b2 and a1 == a2 (the same ciphertext) the TensorSquare form
 * (ElementWise.cu:549-575) reads the two polynomials once. */
int ck_tensor(ck_context *ctx, int num_main, int num_ter, uint32_t *d0,
              uint32_t *d1, uint32_t *d2, const uint32_t *b1,
              const uint32_t *a1, const uint32_t *b2, const uint32_t *a2,
              void *stream);
/* ElementWiseHandler::Permute (ElementWise.cu:966-1002, kernel :370-391):
 * NTT-domain automorphism by 5^rot (rot >= 0) or conjugation (rot == -1);
 * dst != x. */
int ck_permute(ck_context *ctx, int num_main, int num_ter, int num_aux,
               int npoly, uint32_t *const *dst, const uint32_t *const *x,
               int rot, void *stream);
/* Device-to-device copy of `bytes` bytes (DeviceVector's device copies,
 * src/core/DeviceVector.cpp:105-124; ModUp's copy of a group's own limbs,
 * src/core/ModSwitch.cu:467-469, uses the same kernel): 16-byte vectors
 * with non-temporal accesses when dst and src are 16-byte aligned and bytes
 * is a multiple of 16, else hipMemcpyAsync. dst == src is a no-op; ranges
 * that overlap otherwise are rejected. */
int ck_copy(ck_context *ctx, void *dst, const void *src, size_t bytes, void *stream);

/* ---- Context evaluator (include/core/Context.h:101-423) --------------- */
/* Ciphertext polynomials are num_q(level) contiguous limbs; keys are
 * host arrays of dnum device pointers to full-size key blocks. */

/* Context::MultKeyNoModDown (src/core/Context.cpp:609-712): accum (num_q +
 * alpha limbs each) = sum_g ModUp(x)_g (*) key_g, plus P (*) add on the q
 * limbs (add_b / add_a may be NULL). */
int ck_mult_key_no_moddown(ck_context *ctx, int level, uint32_t *accum_b,
                           uint32_t *accum_a, const uint32_t *x,
                           const uint32_t *add_b, const uint32_t *add_a,
                           const uint32_t *const *key_b,
                           const uint32_t *const *key_a, void *stream);
/* Context::MultKey (src/core/Context.cpp:584-607) of a ciphertext (b, a) */
int ck_mult_key(ck_context *ctx, int level, uint32_t *out_b, uint32_t *out_a,
                const uint32_t *b, const uint32_t *a,
                const uint32_t *const *key_b, const uint32_t *const *key_a,
                void *stream);
/* Context::Relinearize / RelinearizeRescale (src/core/Context.cpp:561-565,
 * :714-737) of a 3-poly ciphertext (d0, d1, d2) */
int ck_relinearize(ck_context *ctx, int level, uint32_t *out_b,
                   uint32_t *out_a, const uint32_t *d0, const uint32_t *d1,
                   const uint32_t *d2, const uint32_t *const *key_b,
                   const uint32_t *const *key_a, int rescale, void *stream);
/* Context::HMult (src/core/Context.cpp:862-871): Tensor + Relinearize, with
 * the merged ModDownAndRescale when rescale != 0 (output num_q(level - 1)
 * limbs per poly, else num_q(level)). */
int ck_hmult(ck_context *ctx, int level, uint32_t *out_b, uint32_t *out_a,
             const uint32_t *b1, const uint32_t *a1, const uint32_t *b2,
             const uint32_t *a2, const uint32_t *const *key_b,
             const uint32_t *const *key_a, int rescale, void *stream);
/* ck_hmult without allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. Identical arguments
 * and outputs to ck_hmult; its temporaries live in a static device array of
 * 192 limbs of N = 2^16 words (48 MiB per device, held until the process
 * ends) instead of the context's workspace, so it calls no allocator. Levels
 * whose HMult needs more than that (relinearization workspace above 192
 * limbs) are rejected with CK_ERR_INVALID_ARG, as is a call with
 * per-kernel profiling on. Every call on one device shares that array:
 * calls that overlap on one device (several streams or host threads) race. */
int ck_hmult_noalloc(ck_context *ctx, int level, uint32_t *out_b, uint32_t *out_a,
                     const uint32_t *b1, const uint32_t *a1, const uint32_t *b2,
                     const uint32_t *a2, const uint32_t *const *key_b,
                     const uint32_t *const *key_a, int rescale, void *stream);
/* Context::HRot (src/core/Context.cpp:776-791): MultKey then Permute by
 * 5^rot; rot is reduced mod N/2 (num_slots), rot == 0 copies. */
int ck_hrot(ck_context *ctx, int level, uint32_t *out_b, uint32_t *out_a,
            const uint32_t *b, const uint32_t *a, const uint32_t *const *key_b,
            const uint32_t *const *key_a, int rot, void *stream);
/* Context::HConj (src/core/Context.cpp:793-798) */
int ck_hconj(ck_context *ctx, int level, uint32_t *out_b, uint32_t *out_a,
             const uint32_t *b, const uint32_t *a, const uint32_t *const *key_b,
             const uint32_t *const *key_a, void *stream);
/* Context::Rescale (src/core/Context.cpp:739-774) of npoly (2 or 3)
 * polynomials at `level` -> level - 1 */
int ck_rescale(ck_context *ctx, int level, int npoly, uint32_t *const *dst,
               const uint32_t *const *src, void *stream);

/* Hoisted rotations: HoistHandler::EvaluateBabyStep with the fused key
 * product (src/extension/Hoist.cu:578-691, BSFusedKeyMult :207-327,
 * BSFusedKernel :21-78). One ModUp of a serves n_rot rotations by
 * rots[i] in [1, N/2); key_b / key_a are n_rot x dnum device pointers
 * (row i: the rotation key of rots[i]). moddown != 0: out_*[i] =
 * HRot(ct, rots[i]) (num_q limbs, identical to ck_hrot); moddown == 0:
 * the baby-step output in the Q*P basis (num_q + alpha limbs), as the
 * reference keeps it for the giant step. */
int ck_hrot_hoisted(ck_context *ctx, int level, int n_rot, const int *rots,
                    uint32_t *const *out_b, uint32_t *const *out_a,
                    const uint32_t *b, const uint32_t *a,
                    const uint32_t *const *key_b, const uint32_t *const *key_a,
                    int moddown, void *stream);

/* ---- Batched plans -------------------------------------------------------- */
/* A plan runs one operation over many ciphertexts (any mix of levels) with
 * each step merged into a single launch across them. It has no reference
 * counterpart as an object: it replaces loops of Context::HMult
 * (src/core/Context.cpp:862-871) and of HoistHandler::EvaluateBabyStep
 * (src/extension/Hoist.cu:578-691) over a batch. Creation allocates the
 * plan's workspace and uploads its launch tables; ck_plan_run only launches
 * kernels (no allocation, copy or synchronisation), so it may be captured
 * into a HIP graph. The plan keeps the pointers it was given: buffers must
 * stay allocated until ck_plan_destroy. */
typedef struct ck_plan ck_plan;
/* out_* [i] = HMult(ct1[i], ct2[i]) at levels[i] with one shared key. The
 * items run in chunks one after another (items k, k + chunks, ... form
 * chunk k), each chunk one launch per step: by default as many chunks as
 * keep each chunk's workspace within 512 MiB, so that a step's outputs stay
 * in the last-level cache for the next step. */
int ck_plan_hmult(ck_context *ctx, int n, const int *levels, int rescale,
                  uint32_t *const *out_b, uint32_t *const *out_a,
                  const uint32_t *const *b1, const uint32_t *const *a1,
                  const uint32_t *const *b2, const uint32_t *const *a2,
                  const uint32_t *const *key_b, const uint32_t *const *key_a,
                  ck_plan **out);
/* out_*[i * n_rot + r] = hoisted rotation r of ciphertext i (as
 * ck_hrot_hoisted), one shared set of n_rot rotation keys. */
int ck_plan_hrot_hoisted(ck_context *ctx, int n, const int *levels,
                         const uint32_t *const *b, const uint32_t *const *a,
                         int n_rot, const int *rots,
                         const uint32_t *const *key_b,
                         const uint32_t *const *key_a, uint32_t *const *out_b,
                         uint32_t *const *out_a, int moddown, ck_plan **out);
/* out dst[k] = fast base conversion of src[k] on the FP64 matrix cores,
 * k < n: the reference's ModSwitchMatrixMult (src/core/ModSwitch.cu:23-171)
 * with the PopulateModSwitchConstants tables (:320-366) as a standalone
 * step. From n_src (1..24) source primes s_i to n_dst (1..64) destination
 * primes t_m (global prime indices):
 *   dst[k] limb m = [sum_i v_i (prod_{j != i} s_j) R]_{t_m}  (R = 2^32, canonical)
 * for src[k] = n_src int32 limbs v_i centred in [-(s_i - 1)/2, (s_i - 1)/2]
 * (the INTT-and-constant output of a ModUp or ModDown), exactly the
 * reference's ReduceMontgomery(sum_i v_i table_mi). One launch per run. */
int ck_plan_bconv(ck_context *ctx, const int *src_primes, int n_src,
                  const int *dst_primes, int n_dst, int n, const int32_t *const *src,
                  uint32_t *const *dst, ck_plan **out);
/* dst[k] = ck_ntt (inverse = 0) or ck_intt (inverse = 1) of src[k], k < n,
 * all with one NP (NTTHandler::NTT / INTT over a batch): one launch per pass
 * across the batch. The int8 form's column pass writes a tiled intermediate
 * over dst, so a plan whose outputs overlap its inputs (in place) uses the
 * vector form. */
int ck_plan_ntt(ck_context *ctx, int n, int num_main, int num_ter, int num_aux,
                int inverse, int montgomery_conversion, const uint32_t *const *src,
                uint32_t *const *dst, ck_plan **out);
int ck_plan_run(ck_plan *plan, void *stream);
/* out[8]: items, steps, kernel launches per run, workspace bytes, BaseConv
 * form of an HMult / hoisted plan (0: fused, 1: FP64 matrix cores, 2: int8
 * matrix cores), with
 * CK_OPT_PLAN_AUTOTUNE the measured run time of each of those forms in ns
 * (fused, matrix cores; -1000 when not tuned), and the form of a transform
 * or BaseConv plan (CK_OPT_PLAN_NTT_FORM / CK_OPT_PLAN_BCONV_FORM values) or
 * the key pass of an HMult plan (1: the ModUp's row stages fused with the
 * key inner product, 0: a row pass plus a separate accumulation) */
int ck_plan_info(const ck_plan *plan, long long *out);
/* Autotuning record of a plan built with an autotuned form option: for up
 * to max candidates, forms[3 i .. 3 i + 2] = (form, int8 blocks per
 * workgroup, int8 chunk in limbs) of a transform or BaseConv plan, or
 * (BaseConv form, key pass (as in ck_plan_info), item chunks) of an HMult /
 * hoisted plan, and us[i] = its measured run time in
 * microseconds. Returns the number of candidates (0: not autotuned), or -1
 * on error. No reference counterpart (the reference fixes its launch
 * shapes at compile time, include/core/NTTUtils.cuh:395-437,
 * src/core/ModSwitch.cu:11-19). */
int ck_plan_tuning(const ck_plan *plan, int max, int *forms, double *us);
int ck_plan_destroy(ck_plan *plan);

/* ---- Options ------------------------------------------------------------------ */
/* CK_OPT_BCONV_MFMA: 1 runs every BaseConv (ModUp, ModDown, Rescale) on the
 * FP64 matrix cores as its own step, 2 on the int8 matrix cores (residues
 * and table split into bytes), 0 fuses it into the forward NTT's column
 * pass, -1 (default) uses the int8 matrix cores for batched plans and the
 * fused form for single operations; results are identical. No reference counterpart (the
 * reference's kFuseModDownEpilogue-style compile-time switches,
 * include/core/Parameter.h:13-20, select kernel fusions the same way). */
#define CK_OPT_BCONV_MFMA 1
/* CK_OPT_PLAN_AUTOTUNE: 1 makes ck_plan_hmult / ck_plan_hrot_hoisted build
 * the plan with each BaseConv form (fused, FP64 matrix cores, int8 matrix
 * cores; for ck_plan_hmult each with both key passes, CK_OPT_FUSED_KEY_ROWS
 * 1 and 0, all in one chunk, then the fastest of those in 2, 4 and 8 item
 * chunks), run each (one warm-up + 3 timed runs on a private stream,
 * synchronising the host: plan creation only; the runs write the plan's
 * outputs) and keep the fastest (ck_plan_tuning lists the times); 0
 * (default) uses CK_OPT_BCONV_MFMA and the default chunks. */
#define CK_OPT_PLAN_AUTOTUNE 2
/* CK_OPT_FUSED_KEY_ROWS: 1 (default) runs the last ModUp transform stages
 * (forward row stages) and the key inner product of every key switch
 * (HMult, Relinearize, MultKey, HRot, HConj and their plans) as one pass;
 * 0 runs them as a row pass plus a separate accumulation. Results are
 * identical. */
#define CK_OPT_FUSED_KEY_ROWS 3
/* CK_OPT_PLAN_NTT_FORM: the form of forward transforms in ck_plan_ntt.
 * 0: radix-16 FP64 butterflies on the vector unit (NTTPhase1 / NTTPhase2,
 * src/core/NTT.cu:231-454); 1 (default, measured faster): the 4-step
 * 256 x 256 transform as exact int8 matrix products (residues split into
 * four bytes, int32 accumulation, exact recombination); -1: autotune, i.e. build the vector
 * form and the int8 form for every (CK_OPT_NTT8_BLOCKS, CK_OPT_NTT8_CHUNK)
 * candidate, time each and keep the fastest (ck_plan_tuning lists them;
 * outputs must not alias inputs). Inverse transforms always use the vector
 * form. Results are identical. */
#define CK_OPT_PLAN_NTT_FORM 4
/* CK_OPT_PLAN_BCONV_FORM: the form of ck_plan_bconv. 0: FP64 matrix cores
 * (16-bit table halves); 1 (default, measured faster): int8 matrix cores
 * (residues and table split into bytes); 2: vector unit (exact FP64 FMAs);
 * -1: autotune over the three. Results are identical. */
#define CK_OPT_PLAN_BCONV_FORM 5
/* CK_OPT_NTT8_BLOCKS: 16-column (16-row) blocks of one limb per workgroup
 * of the int8 transform's column (row) pass: 1, 2, 4, 8 or 16 (default). */
#define CK_OPT_NTT8_BLOCKS 6
/* CK_OPT_NTT8_CHUNK: limbs per chunk of an int8 transform plan, whose
 * column and row passes run chunk by chunk so that a chunk's intermediate
 * can stay in the last-level cache; 0 (default): one chunk. */
#define CK_OPT_NTT8_CHUNK 7
int ck_context_set_option(ck_context *ctx, int option, int value);

/* ---- Per-kernel timing ---------------------------------------------------- */
/* Replaces the wall-clock profiling macros of the reference's test harness
 * (unittest/Testbed.h:24-47) with device-side timing: while enabled, every
 * kernel the context launches is bracketed by HIP events on the caller's
 * stream (enabling or disabling clears earlier records). */
int ck_profile_enable(ck_context *ctx, int on);
/* Waits for the recorded events and writes, per kernel name, the summed
 * duration and the launch count; names is max_entries x 64 bytes. Returns
 * the number of entries (or -1 on error). */
int ck_profile_report(ck_context *ctx, int max_entries, char *names,
                      double *total_ms, long long *calls);

#ifdef __cplusplus
}
#endif
#endif /* CHEDDAR_AMD_H_ */
