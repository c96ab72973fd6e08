// Batched plans: many ciphertexts (at any mix of levels) through one
// operation, each step merged across them into a single launch.
//
// The reference runs HMult / hoisted rotations one ciphertext at a time
// (src/core/Context.cpp:862-871, src/extension/Hoist.cu:578-691). Here a
// plan is built once: the same op code that drives the single-ciphertext
// path (ck_ops.cpp) is replayed per ciphertext into a PlanSink, which
// appends each step's per-limb jobs to that step's table; the tables and a
// workspace for every ciphertext are uploaded at creation. Running a plan
// is then one launch per step (table-driven kernels), with no allocation,
// copy or synchronisation, so it can be captured into a HIP graph.
#include "ck_plan.h"

#include <algorithm>
#include <cstring>

namespace ck {

namespace {

constexpr size_t kLimb = 65536;

template <typename T>
int upload_vec(Plan &p, T **dst, const std::vector<T> &v) {
  *dst = nullptr;
  if (v.empty()) return kOk;
  int rc = check_hip(hipMalloc((void **)dst, v.size() * sizeof(T)),
                     "hipMalloc(plan table)");
  if (rc) return rc;
  p.allocations.push_back(*dst);
  return check_hip(hipMemcpy(*dst, v.data(), v.size() * sizeof(T),
                             hipMemcpyHostToDevice),
                   "hipMemcpy(plan table)");
}

}  // namespace

PlanStage &PlanSink::stage(PlanStage::Kind kind) {
  if (cursor_ == plan_.stages.size()) {
    plan_.stages.emplace_back();
    plan_.stages.back().kind = kind;
  }
  PlanStage &s = plan_.stages[cursor_++];
  if (s.kind != kind) mismatch_ = true;
  return s;
}

int PlanSink::pass(bool forward, int flags, const std::vector<PassJob> &jobs) {
  PlanStage &s = stage(forward ? PlanStage::kFwd : PlanStage::kInv);
  s.flags_any |= flags;
  for (const PassJob &j : jobs) {
    LimbJob J;
    std::memset(&J, 0, sizeof(J));
    J.src = j.src + (size_t)j.src_limb * kLimb;
    J.dst = j.dst + (size_t)j.dst_limb * kLimb;
    J.src2 = (j.src2 && j.src2_limb >= 0) ? j.src2 + (size_t)j.src2_limb * kLimb
                                          : nullptr;
    J.bsrc = j.bsrc;
    J.btab = j.btab ? j.btab + j.btab_off : nullptr;
    J.bs = (uint16_t)j.bs;
    J.c1 = j.c1;
    J.c2 = j.c2;
    J.c3 = j.c3;
    J.prime = (uint16_t)j.prime;
    J.flags = flags;
    s.max_bs = std::max(s.max_bs, j.bs);
    s.jobs.push_back(J);
  }
  return kOk;
}

int PlanSink::copy(uint32_t *, const uint32_t *, size_t) {
  set_error("device copies are not part of batched plans");
  return kErrUnsupported;
}

int PlanSink::bconv(const std::vector<BconvInst> &insts) {
  if (plan_.bconv_form == 2) {  // int8 matrix cores
    PlanStage &s = stage(PlanStage::kBconvI8);
    for (const BconvInst &I : insts) {
      BconvI8Inst J;
      if (!to_i8_inst(*plan_.ctx, I, &J)) return kErrHip;
      s.bi.push_back(J);
    }
    return kOk;
  }
  PlanStage &s = stage(PlanStage::kBconv);
  s.bc.insert(s.bc.end(), insts.begin(), insts.end());
  return kOk;
}

int PlanSink::key_accum(const EwNp &np, const KeyAccumArgs &A) {
  PlanStage &s = stage(PlanStage::kKeyAccum);
  const uint16_t item = (uint16_t)s.ka.size();
  s.ka.push_back(KeyAccumItem{np, A});
  for (int y = 0; y < np.num_total; y++) s.ew.push_back(EwJob{item, (uint16_t)y});
  return kOk;
}

int PlanSink::key_rows(const EwNp &np, const KeyAccumArgs &A) {
  PlanStage &s = stage(PlanStage::kKeyRows);
  const uint16_t item = (uint16_t)s.ka.size();
  s.ka.push_back(KeyAccumItem{np, A});
  for (int y = 0; y < np.num_total; y++) s.ew.push_back(EwJob{item, (uint16_t)y});
  return kOk;
}

int PlanSink::bs_fused(const EwNp &np, const BsFusedArgs &A) {
  PlanStage &s = stage(PlanStage::kBsFused);
  const uint16_t item = (uint16_t)s.bs.size();
  s.bs.push_back(BsFusedItem{np, A});
  for (int y = 0; y < np.num_total; y++) s.ew.push_back(EwJob{item, (uint16_t)y});
  return kOk;
}

int PlanSink::permute(const EwNp &, const EwOperand &, const EwOperand &,
                      uint32_t, int) {
  set_error("permutations are not part of batched plans");
  return kErrUnsupported;
}

namespace {

int ew_prime(const EwNp &np, int y) {
  return y < np.num_q ? np.q_prime0 + y : np.aux_prime0 + (y - np.num_q);
}

// Jobs of one step are independent, so their order is free: grouping them by
// prime lets consecutive workgroups share the twiddle table (transforms) or
// the key limb (key products, whose keys every item shares) in cache instead
// of each item fetching it from HBM.
// A BaseConv-input pass is the exception: its output limbs of one
// polynomial re-read the same source tiles, so they stay together.
void order_by_prime(PlanStage &s) {
  if (!(s.flags_any & kNttBconvInput))
    std::stable_sort(s.jobs.begin(), s.jobs.end(),
                     [](const LimbJob &a, const LimbJob &b) { return a.prime < b.prime; });
  const bool ka = s.kind == PlanStage::kKeyAccum || s.kind == PlanStage::kKeyRows;
  auto prime_of = [&](const EwJob &j) {
    return ka ? ew_prime(s.ka[j.item].np, j.y) : ew_prime(s.bs[j.item].np, j.y);
  };
  if (ka || s.kind == PlanStage::kBsFused)
    std::stable_sort(s.ew.begin(), s.ew.end(), [&](const EwJob &a, const EwJob &b) {
      return prime_of(a) < prime_of(b);
    });
}

}  // namespace

int plan_finish(Plan &p, bool mismatch) {
  if (mismatch) {
    set_error("plan items do not share one step sequence");
    return kErrUnsupported;
  }
  for (PlanStage &s : p.stages) {
    order_by_prime(s);
    if (s.ka.size() > 65535 || s.bs.size() > 65535 || s.bc.size() > 65535 ||
        s.bi.size() > 65535) {
      set_error("too many items in one plan step");
      return kErrUnsupported;
    }
    int rc = upload_vec(p, &s.d_jobs, s.jobs);
    if (!rc) rc = upload_vec(p, &s.d_ka, s.ka);
    if (!rc) rc = upload_vec(p, &s.d_bs, s.bs);
    if (!rc) rc = upload_vec(p, &s.d_ew, s.ew);
    if (!rc) rc = upload_vec(p, &s.d_bc, s.bc);
    if (!rc) rc = upload_vec(p, &s.d_n8, s.n8);
    if (!rc) rc = upload_vec(p, &s.d_bi, s.bi);
    if (rc) return rc;
  }
  return kOk;
}

int plan_upload(Plan &p, double **dst, const std::vector<double> &v) {
  return upload_vec(p, dst, v);
}

int plan_upload_i64(Plan &p, int64_t **dst, const std::vector<int64_t> &v) {
  return upload_vec(p, dst, v);
}

void plan_add_ntt8(Plan &p, std::vector<Ntt8Job> jobs, int blocks, int chunk) {
  std::stable_sort(jobs.begin(), jobs.end(),
                   [](const Ntt8Job &a, const Ntt8Job &b) { return a.prime < b.prime; });
  const size_t n = jobs.size();
  const size_t step = chunk > 0 ? (size_t)chunk : n;
  for (size_t b = 0; b < n; b += step) {
    const size_t e = std::min(n, b + step);
    for (int pass = 0; pass < 2; pass++) {
      p.stages.emplace_back();
      PlanStage &s = p.stages.back();
      s.kind = pass ? PlanStage::kNtt8Rows : PlanStage::kNtt8Cols;
      s.blocks = blocks;
      s.n8.assign(jobs.begin() + b, jobs.begin() + e);
    }
  }
}

int plan_alloc_workspace(Plan &p, size_t limbs, uint32_t **ws) {
  *ws = nullptr;
  if (limbs == 0) return kOk;
  int rc = check_hip(hipMalloc((void **)ws, limbs * kLimb * sizeof(uint32_t)),
                     "hipMalloc(plan workspace)");
  if (rc) return rc;
  p.allocations.push_back(*ws);
  p.workspace_bytes += limbs * kLimb * sizeof(uint32_t);
  return kOk;
}

int plan_run(const Plan &p, hipStream_t st) {
  const Context &c = *p.ctx;
  for (const PlanStage &s : p.stages) {
    switch (s.kind) {
      case PlanStage::kFwd:
        launch_ntt_fwd_table(c, s.d_jobs, (int)s.jobs.size(), s.flags_any, s.max_bs, st);
        break;
      case PlanStage::kInv:
        launch_intt_table(c, s.d_jobs, (int)s.jobs.size(), s.flags_any, st);
        break;
      case PlanStage::kKeyAccum:
        launch_key_accum_table(c, s.d_ka, s.d_ew, (int)s.ew.size(), st);
        break;
      case PlanStage::kKeyRows:
        launch_key_rows_table(c, s.d_ka, s.d_ew, (int)s.ew.size(), st);
        break;
      case PlanStage::kBsFused:
        launch_bs_fused_table(c, s.d_bs, s.d_ew, (int)s.ew.size(), st);
        break;
      case PlanStage::kNtt8Cols:
        launch_ntt8_cols(c, p.i8, s.d_n8, (int)s.n8.size(), s.blocks, st);
        break;
      case PlanStage::kNtt8Rows:
        launch_ntt8_rows(c, p.i8, s.d_n8, (int)s.n8.size(), s.blocks, st);
        break;
      case PlanStage::kBconvI8:
      case PlanStage::kBconvValu: {
        int dsts = 0, max_s = 1, max_d = 1;
        for (const BconvI8Inst &I : s.bi) {
          dsts += I.D;
          max_s = std::max(max_s, I.S);
          max_d = std::max(max_d, I.D);
        }
        if (s.kind == PlanStage::kBconvI8)
          launch_bconv_i8_table(c, s.d_bi, (int)s.bi.size(), dsts, max_s, max_d, st);
        else
          launch_bconv_valu_table(c, s.d_bi, (int)s.bi.size(), dsts, max_s, st);
        break;
      }
      case PlanStage::kBconv: {
        int dsts = 0, max_s = 1, max_d = 1;
        for (const BconvInst &I : s.bc) {
          dsts += I.D;
          max_s = std::max(max_s, I.S);
          max_d = std::max(max_d, I.D);
        }
        launch_bconv_mfma_table(c, s.d_bc, (int)s.bc.size(), dsts, max_s, max_d, st);
        break;
      }
    }
  }
  return check_hip(hipGetLastError(), "plan run");
}

void plan_destroy(Plan *p) {
  if (!p) return;
  for (void *a : p->allocations) (void)hipFree(a);
  delete p;
}

double plan_time_us(const Plan &p, int reps) {
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double us = -1.0;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess &&
      hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
      hipDeviceSynchronize() == hipSuccess && plan_run(p, st) == kOk &&
      hipEventRecord(e0, st) == hipSuccess) {
    bool ok = true;
    for (int r = 0; r < reps && ok; r++) ok = plan_run(p, st) == kOk;
    float ms = 0.f;
    if (ok && hipEventRecord(e1, st) == hipSuccess &&
        hipEventSynchronize(e1) == hipSuccess &&
        hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
      us = 1e3 * ms / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) {
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
  }
  return us;
}

int plan_launches(const Plan &p) {
  int n = 0;
  for (const PlanStage &s : p.stages)
    n += ((s.kind == PlanStage::kFwd && !(s.flags_any & kNttColsOnly)) ||
          s.kind == PlanStage::kInv)
             ? 2
             : 1;
  return n;
}

}  // namespace ck
