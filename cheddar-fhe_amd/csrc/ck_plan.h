// Batched plans (see ck_plan.cpp).
#pragma once

#include "ck_i8.h"
#include "ck_ops.h"

namespace ck {

// One step of a plan: a transform pass over every job's limb, or one
// key-accumulation kernel over every item's limbs.
struct PlanStage {
  enum Kind {
    kFwd, kInv, kKeyAccum, kKeyRows, kBsFused, kBconv,
    kNtt8Cols, kNtt8Rows,   // int8 matrix-core transform passes
    kBconvI8, kBconvValu    // BaseConv on the int8 matrix cores / vector unit
  } kind = kFwd;
  int flags_any = 0;
  int max_bs = 0;  // largest BaseConv source count among the jobs
  std::vector<LimbJob> jobs;
  std::vector<KeyAccumItem> ka;
  std::vector<BsFusedItem> bs;
  std::vector<EwJob> ew;
  std::vector<BconvInst> bc;
  std::vector<Ntt8Job> n8;      // kNtt8Cols / kNtt8Rows
  std::vector<BconvI8Inst> bi;  // kBconvI8 / kBconvValu
  int blocks = 16;              // int8 passes: 16-column / 16-row blocks per workgroup
  LimbJob *d_jobs = nullptr;
  BconvInst *d_bc = nullptr;
  Ntt8Job *d_n8 = nullptr;
  BconvI8Inst *d_bi = nullptr;
  KeyAccumItem *d_ka = nullptr;
  BsFusedItem *d_bs = nullptr;
  EwJob *d_ew = nullptr;
};

struct Plan {
  const Context *ctx = nullptr;
  std::vector<PlanStage> stages;
  std::vector<void *> allocations;
  size_t workspace_bytes = 0;
  int items = 0;
  bool bconv_mfma = true;  // BaseConv steps on the matrix cores
  int bconv_form = 2;       // 0 fused, 1 FP64 matrix cores, 2 int8 matrix cores
  int key_rows = -1;        // key pass: 1 fused row stages, 0 separate, -1 the context's
  int chunks = 1;           // item chunks, run one after another
  // autotuning record: measured run time of each BaseConv form (us; < 0:
  // not measured): fused, FP64 matrix cores, int8 matrix cores
  double tuned_us[3] = {-1.0, -1.0, -1.0};
  // the form of a transform or BaseConv plan (ck_plan_info) and every
  // candidate that autotuning timed: form, blocks, chunk, run time (us)
  int form = 0;
  struct Candidate {
    int form, blocks, chunk;
    double us;
  };
  std::vector<Candidate> tuning;
  I8NttTabs i8{};
};

// Records the ops of one item after another; begin_item() rewinds to the
// first step so that step k of every item lands in the same table.
class PlanSink : public Sink {
 public:
  explicit PlanSink(Plan &p) : plan_(p) {}
  bool mfma_bconv() const override { return plan_.bconv_mfma; }
  bool fused_key_rows(const Context &c) const override {
    return plan_.key_rows >= 0 ? plan_.key_rows != 0 : c.fuse_key_rows;
  }
  void begin_item() { cursor_ = base_; }
  // the items recorded after this run as a chunk of their own: their steps
  // are appended after the previous chunk's, so a run issues the chunks one
  // after another
  void begin_chunk() { base_ = plan_.stages.size(); }
  bool mismatch() const { return mismatch_; }
  int pass(bool forward, int flags, const std::vector<PassJob> &jobs) override;
  int copy(uint32_t *dst, const uint32_t *src, size_t words) override;
  int bconv(const std::vector<BconvInst> &insts) override;
  int key_accum(const EwNp &np, const KeyAccumArgs &A) override;
  int key_rows(const EwNp &np, const KeyAccumArgs &A) override;
  int bs_fused(const EwNp &np, const BsFusedArgs &A) override;
  int permute(const EwNp &np, const EwOperand &x, const EwOperand &out,
              uint32_t g, int npoly) override;

 private:
  PlanStage &stage(PlanStage::Kind kind);
  Plan &plan_;
  size_t cursor_ = 0;
  size_t base_ = 0;
  bool mismatch_ = false;
};

int plan_alloc_workspace(Plan &p, size_t limbs, uint32_t **ws);
// a constant table the plan owns (freed by plan_destroy)
int plan_upload(Plan &p, double **dst, const std::vector<double> &v);
int plan_finish(Plan &p, bool mismatch);
// Appends the int8 forward transform of `jobs` to p: column pass then row
// pass per chunk of `chunk` limbs (0: all jobs in one chunk), so a chunk's
// intermediate can stay in the last-level cache between its two passes;
// jobs are ordered by prime first.
void plan_add_ntt8(Plan &p, std::vector<Ntt8Job> jobs, int blocks, int chunk);
// a constant int table the plan owns (freed by plan_destroy)
int plan_upload_i64(Plan &p, int64_t **dst, const std::vector<int64_t> &v);
int plan_run(const Plan &p, hipStream_t st);
void plan_destroy(Plan *p);
int plan_launches(const Plan &p);
// mean run time of a plan over `reps` runs after one warm-up, on a private
// stream (host-synchronising; for plan creation only)
double plan_time_us(const Plan &p, int reps);

}  // namespace ck
