// Host orchestration entry points used by the C ABI (see ck_ops.cpp).
#pragma once

#include "ck_elem.h"

namespace ck {

int op_ntt(const Context &c, uint32_t *dst, const uint32_t *src, int nm,
           int nt, int na, int src_gap, bool mont_conv, hipStream_t st);
int op_intt(const Context &c, uint32_t *dst, const uint32_t *src, int nm,
            int nt, int na, int src_gap, bool mont_conv, hipStream_t st);
int op_intt_const(const Context &c, uint32_t *dst, const uint32_t *src,
                  int nm, int nt, int na, int src_gap,
                  const uint32_t *consts_mont, bool centred, hipStream_t st);
int op_modup(const Context &c, int level, uint32_t *dst, const uint32_t *src,
             uint32_t *v, hipStream_t st, bool copy_own = true,
             const uint32_t *prod2 = nullptr);
const ModDownPlan *moddown_plan(const Context &c, int level, int kind);
int op_moddown(const Context &c, int level, int kind, int npoly,
               uint32_t *const *dst, const uint32_t *const *src, uint32_t *v,
               hipStream_t st);
size_t mult_key_ws_limbs(const Context &c, int level);
int op_mult_key_no_moddown(const Context &c, int level, uint32_t *acc_b,
                           uint32_t *acc_a, const uint32_t *x,
                           const uint32_t *add_b, const uint32_t *add_a,
                           const uint32_t *const *key_b,
                           const uint32_t *const *key_a, uint32_t *ws,
                           hipStream_t st,
                           const uint32_t *const *tensor = nullptr);
int op_relinearize(const Context &c, int level, uint32_t *out_b,
                   uint32_t *out_a, const uint32_t *d0, const uint32_t *d1,
                   const uint32_t *d2, const uint32_t *const *key_b,
                   const uint32_t *const *key_a, bool rescale, uint32_t *ws,
                   hipStream_t st);
int op_hmult(const Context &c, int level, uint32_t *out_b, uint32_t *out_a,
             const uint32_t *b1, const uint32_t *a1, const uint32_t *b2,
             const uint32_t *a2, const uint32_t *const *key_b,
             const uint32_t *const *key_a, bool rescale, hipStream_t st);
int op_mult_key(const Context &c, int level, uint32_t *out_b, uint32_t *out_a,
                const uint32_t *b, const uint32_t *a,
                const uint32_t *const *key_b, const uint32_t *const *key_a,
                uint32_t *ws, hipStream_t st);
int op_hrot_galois(const Context &c, int level, uint32_t *out_b,
                   uint32_t *out_a, const uint32_t *b, const uint32_t *a,
                   const uint32_t *const *key_b, const uint32_t *const *key_a,
                   uint32_t galois, hipStream_t st);

}  // namespace ck
