// Host orchestration of the key-switching path (see ck_ops.cpp).
//
// Every operation is written once against a Sink: the sequence of launches
// it needs (transform passes, key accumulations, ...). An ImmediateSink
// launches them at once on a stream (the single-operation C ABI); a plan
// sink (ck_plan.cpp) records the same sequence for many ciphertexts and
// merges each step across them into one launch.
#pragma once

#include "ck_bconv.h"
#include "ck_elem.h"

namespace ck {

// One limb of a transform pass in base + limb form (a limb index is taken
// in units of N words from its base pointer).
struct PassJob {
  const uint32_t *src = nullptr;
  int src_limb = 0;
  uint32_t *dst = nullptr;
  int dst_limb = 0;
  const uint32_t *src2 = nullptr;  // product operand / kept limb
  int src2_limb = -1;              // -1: none
  const int32_t *bsrc = nullptr;   // BaseConv sources (bs contiguous limbs)
  const double *btab = nullptr;    // BaseConv table of the destination set
  int bs = 0;
  int btab_off = 0;                // row of this limb in btab (doubles)
  int prime = 0;
  uint32_t c1 = 0, c2 = 0, c3 = 0;
};

class Sink {
 public:
  virtual ~Sink() = default;
  // forward (column + row pass) or inverse (row + column pass) transform
  virtual int pass(bool forward, int flags, const std::vector<PassJob> &jobs) = 0;
  virtual int copy(uint32_t *dst, const uint32_t *src, size_t words) = 0;
  // BaseConv instances on the matrix cores (outputs: lazy int32 limbs)
  virtual int bconv(const std::vector<BconvInst> &insts) = 0;
  // whether ops should emit bconv() steps rather than fused BaseConv passes
  virtual bool mfma_bconv() const = 0;
  // key switches: the ModUp's forward row stages fused with the key inner
  // product (key_rows) or a row pass plus key_accum
  virtual bool fused_key_rows(const Context &c) const { return c.fuse_key_rows; }
  virtual int key_accum(const EwNp &np, const KeyAccumArgs &A) = 0;
  // forward row stages of a column-pass-only ModUp fused with key_accum
  virtual int key_rows(const EwNp &np, const KeyAccumArgs &A) = 0;
  virtual int bs_fused(const EwNp &np, const BsFusedArgs &A) = 0;
  virtual int permute(const EwNp &np, const EwOperand &x, const EwOperand &out,
                      uint32_t g, int npoly) = 0;
};

// Launches everything at once on one stream; passes are packed into
// kernel-argument plans of at most kMaxLaunchLimbs limbs.
class ImmediateSink : public Sink {
 public:
  ImmediateSink(const Context &c, hipStream_t st) : c_(c), st_(st) {}
  int pass(bool forward, int flags, const std::vector<PassJob> &jobs) override;
  int copy(uint32_t *dst, const uint32_t *src, size_t words) override;
  int bconv(const std::vector<BconvInst> &insts) override;
  bool mfma_bconv() const override { return c_.bconv_mfma == 1 || c_.bconv_mfma == 2; }
  int key_accum(const EwNp &np, const KeyAccumArgs &A) override;
  int key_rows(const EwNp &np, const KeyAccumArgs &A) override;
  int bs_fused(const EwNp &np, const BsFusedArgs &A) override;
  int permute(const EwNp &np, const EwOperand &x, const EwOperand &out,
              uint32_t g, int npoly) override;

 private:
  const Context &c_;
  hipStream_t st_;
};

int op_ntt(const Context &c, uint32_t *dst, const uint32_t *src, int nm,
           int nt, int na, int src_gap, bool mont_conv, Sink &s);
int op_intt(const Context &c, uint32_t *dst, const uint32_t *src, int nm,
            int nt, int na, int src_gap, bool mont_conv, Sink &s);
int op_intt_const(const Context &c, uint32_t *dst, const uint32_t *src,
                  int nm, int nt, int na, int src_gap,
                  const uint32_t *consts_mont, bool centred, Sink &s);
// cols_only: the forward transforms of the outputs stop after the column
// stages (lazy int32 in dst), for a key_rows pass to finish
int op_modup(const Context &c, int level, uint32_t *dst, const uint32_t *src,
             uint32_t *v, Sink &s, bool copy_own = true,
             const uint32_t *prod2 = nullptr, bool cols_only = false);
const ModDownPlan *moddown_plan(const Context &c, int level, int kind);
int op_moddown(const Context &c, int level, int kind, int npoly,
               uint32_t *const *dst, const uint32_t *const *src, uint32_t *v,
               Sink &s);
size_t mult_key_ws_limbs(const Context &c, int level);
int op_mult_key_no_moddown(const Context &c, int level, uint32_t *acc_b,
                           uint32_t *acc_a, const uint32_t *x,
                           const uint32_t *add_b, const uint32_t *add_a,
                           const uint32_t *const *key_b,
                           const uint32_t *const *key_a, uint32_t *ws,
                           Sink &s, const uint32_t *const *tensor = nullptr);
size_t relin_ws_limbs(const Context &c, int level);
int op_relinearize(const Context &c, int level, uint32_t *out_b,
                   uint32_t *out_a, const uint32_t *d0, const uint32_t *d1,
                   const uint32_t *d2, const uint32_t *const *key_b,
                   const uint32_t *const *key_a, bool rescale, uint32_t *ws,
                   Sink &s);
int op_hmult(const Context &c, int level, uint32_t *out_b, uint32_t *out_a,
             const uint32_t *b1, const uint32_t *a1, const uint32_t *b2,
             const uint32_t *a2, const uint32_t *const *key_b,
             const uint32_t *const *key_a, bool rescale, uint32_t *ws,
             Sink &s);
int op_mult_key(const Context &c, int level, uint32_t *out_b, uint32_t *out_a,
                const uint32_t *b, const uint32_t *a,
                const uint32_t *const *key_b, const uint32_t *const *key_a,
                uint32_t *ws, Sink &s);
int op_hrot_galois(const Context &c, int level, uint32_t *out_b,
                   uint32_t *out_a, const uint32_t *b, const uint32_t *a,
                   const uint32_t *const *key_b, const uint32_t *const *key_a,
                   uint32_t galois, uint32_t *ws, Sink &s);

// Hoisted rotations (HoistHandler::EvaluateBabyStep, Hoist.cu:578-691):
// one ModUp of a shared by n_rot rotations. key_b / key_a hold n_rot rows of
// dnum key blocks; rots[i] in [1, N/2). With moddown, out_*[i] is
// HRot(ct, rots[i]) (num_q limbs); without, the Q*P-basis baby-step output
// (num_q + alpha limbs). The workspace (hoisted_ws_limbs) holds the
// accumulators of chunk_rot rotations at a time.
size_t hoisted_ws_limbs(const Context &c, int level, int chunk_rot,
                        bool moddown);
int op_hoisted_rotations(const Context &c, int level, int n_rot,
                         const int *rots, uint32_t *const *out_b,
                         uint32_t *const *out_a, const uint32_t *b,
                         const uint32_t *a, const uint32_t *const *key_b,
                         const uint32_t *const *key_a, bool moddown,
                         int chunk_rot, uint32_t *ws, Sink &s);

}  // namespace ck
