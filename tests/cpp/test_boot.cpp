// Bootstrapping through the C++ host facade on the GPU (the reference's
// unittest/Bootstrapping.cpp: encrypt a random message at level 0,
// bootstrap, decrypt and compare within 1e-3), with stage checks first:
// - EvalMod on values (I + r) / K against the plain evaluation of the same
//   schedule and against sin(2 pi r) / (2 pi);
// - the first CoeffToSlot and SlotToCoeff phases against the plain
//   baby-step / giant-step product of their plaintext maps;
// - the full bootstrap: precision bits = -log2(max error) and the mean
//   error, and the wall time per bootstrap (after one warm-up);
// - four bootstraps captured into one HIP graph: three replays, with eager
//   work and synchronisation between them, each output equal word for word
//   to the eager bootstrap of the same input;
// - 2^14 slots (Trace + doubled transforms) and the imaginary-removing
//   variant at 2^13 slots (the output is the real part of the message).
//
//   test_boot <bootparam.json> [timed_runs] [sse]
//
// By default the main secret has the sparse Hamming weight of the parameter
// file (no sparse-secret encapsulation); with "sse" the file's dense weight
// is kept and ModUpToMax switches to the sparse secret and back.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "cheddar/ParameterJson.h"
#include "cheddar/UserInterface.h"
#include "cheddar/extension/BootContext.h"
#include "cheddar/extension/BootGraph.h"
#include "ckks_oracle.h"

using namespace cheddar;
using word = uint32_t;
using Ct = Ciphertext<word>;
using Pt = Plaintext<word>;
using Msg = std::vector<Complex>;
using Clock = std::chrono::steady_clock;

namespace {

int g_fail = 0;

double Seconds(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

void Report(const std::string &name, const Msg &want, const Msg &got, double tol) {
  double worst = 0, sum = 0;
  bool ok = want.size() == got.size() && !want.empty();
  for (size_t i = 0; ok && i < want.size(); i++) {
    const double e = std::abs(want[i] - got[i]);
    if (!(e <= worst)) worst = e;
    sum += e;
  }
  ok = ok && worst < tol;
  std::printf("%s %s: max error %.3e (%.2f bits), mean error %.3e (%.2f bits)\n",
              ok ? "PASS" : "FAIL", name.c_str(), worst, -std::log2(worst),
              sum / want.size(), -std::log2(sum / want.size()));
  std::fflush(stdout);
  if (!ok) g_fail++;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <bootparam.json> [timed_runs]\n", argv[0]);
    return 2;
  }
  const int runs = argc > 2 ? std::atoi(argv[2]) : 2;
  const bool sse = argc > 3 && std::string(argv[3]) == "sse";
  std::ifstream f(argv[1]);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  auto param = ParameterFromJsonText(text);
  int ncts = 0, nstc = 0;
  if (!BootLevelsFromJsonText(text, &ncts, &nstc)) {
    std::fprintf(stderr, "%s does not enable bootstrapping\n", argv[1]);
    return 2;
  }
  if (!sse) param->SetDenseHammingWeight(param->GetSparseHammingWeight());
  const Parameter<word> &P = *param;
  const int slots = P.degree_ / 2;
  std::printf("secret: dense Hamming weight %d, sparse %d, encapsulation %s\n",
              P.GetDenseHammingWeight(), P.GetSparseHammingWeight(),
              P.IsUsingSparseSecretEncapsulation() ? "on" : "off");

  auto t0 = Clock::now();
  auto ctx = BootContext<word>::Create(P, BootParameter(P.max_level_, ncts, nstc));
  UserInterface<word> ui(ctx, 0x4348454444415204ull);
  ctx->PrepareEvalMod();
  ctx->PrepareEvalSpecialFFT(slots);
  EvkRequest req;
  ctx->AddRequiredRotations(req, slots);
  ctx->AddRequiredRotations(req, slots, true);  // the min-KS strides
  ui.PrepareRotationKey(req);
  (void)hipDeviceSynchronize();
  std::printf("setup: %.1f s (%zu rotation keys; levels: CtS %d, EvalMod %d, StC %d, end %d)\n",
              Seconds(t0), req.size(), ctx->boot_param_.GetCtSStartLevel(),
              ctx->boot_param_.GetEvalModStartLevel(), ctx->boot_param_.GetStCStartLevel(),
              ctx->boot_param_.GetEndLevel());
  std::fflush(stdout);
  const EvkMap<word> &keys = ui.GetEvkMap();
  Random rng(0x4348454444415203ull);

  auto encrypt = [&](Ct &ct, const Msg &m, int level, double scale) {
    Pt pt;
    ctx->encoder_.Encode(pt, level, scale, m);
    ui.Encrypt(ct, pt);
  };
  auto decrypt = [&](const Ct &ct) {
    Pt pt;
    ui.Decrypt(pt, ct);
    Msg m;
    ctx->encoder_.Decode(m, pt);
    return m;
  };

  try {
    // ModUpToMax (BootContext.cpp:186-268) word for word against the oracle:
    // a uniform level -1 ciphertext, with the facade's own DtS / StD keys
    // when encapsulating
    std::vector<int> cfg;
    for (auto &c : P.level_config_) {
      cfg.push_back(c.first);
      cfg.push_back(c.second);
    }
    or_set_num_threads(8);
    or_param *op = or_param_create(P.log_degree_, P.main_primes_.data(), (int)P.main_primes_.size(),
                                   P.ter_primes_.data(), (int)P.ter_primes_.size(),
                                   P.aux_primes_.data(), (int)P.aux_primes_.size(), cfg.data(),
                                   (int)P.level_config_.size());
    if (!op) throw Error("oracle rejected the parameters");
    const NPInfo base = P.LevelToNP(-1);
    const size_t bw = (size_t)base.GetNumQ() * P.degree_;
    std::vector<word> hb(bw), ha(bw);
    or_fill_uniform(op, base.num_main_, base.num_ter_, 0, 0x4348454444415205ull, hb.data());
    or_fill_uniform(op, base.num_main_, base.num_ter_, 0, 0x4348454444415206ull, ha.data());
    Ct ct, out;
    ct.ModifyNP(base);
    ct.bx_.CopyFromHost(hb);
    ct.ax_.CopyFromHost(ha);
    ctx->ModUpToMax(out, ct, keys);
    const std::vector<word> gb = out.bx_.CopyToHost(), ga = out.ax_.CopyToHost();
    const size_t tw = (size_t)P.LevelToNP(P.max_level_).GetNumQ() * P.degree_;
    std::vector<word> wb(tw), wa(tw);
    int rc;
    if (P.IsUsingSparseSecretEncapsulation()) {
      const EvaluationKey<word> &dts = keys.GetDenseToSparseKey();
      const EvaluationKey<word> &sd = keys.GetSparseToDenseKey();
      const std::vector<word> db = dts.bx_[0].CopyToHost(), da = dts.ax_[0].CopyToHost();
      const std::vector<word> sb = sd.bx_[0].CopyToHost(), sa = sd.ax_[0].CopyToHost();
      rc = or_mod_up_to_max(op, base.num_main_, base.num_ter_, 1, hb.data(), ha.data(),
                            db.data(), da.data(), sb.data(), sa.data(), wb.data(), wa.data());
    } else {
      rc = or_mod_up_to_max(op, base.num_main_, base.num_ter_, 0, hb.data(), ha.data(), nullptr,
                            nullptr, nullptr, nullptr, wb.data(), wa.data());
    }
    or_param_destroy(op);
    size_t diff = 0;
    for (size_t i = 0; i < tw; i++) diff += (gb.size() == tw && gb[i] != wb[i]) + (ga.size() == tw && ga[i] != wa[i]);
    const bool same = rc == 0 && gb.size() == tw && ga.size() == tw && diff == 0;
    std::printf("%s ModUpToMax%s vs the oracle, word for word: %zu of %zu words differ\n",
                same ? "PASS" : "FAIL",
                P.IsUsingSparseSecretEncapsulation() ? " (dense -> sparse -> dense)" : "", diff,
                2 * tw);
    std::fflush(stdout);
    if (!same) g_fail++;
  } catch (const std::exception &e) {
    std::printf("FAIL ModUpToMax: %s\n", e.what());
    g_fail++;
  }

  try {
    // EvalMod: t = (I + r) / K -> ~ sin(2 pi r) / (2 pi)
    const BootParameter &bp = ctx->boot_param_;
    const int K = (1 << bp.num_double_angle_) * bp.initial_K_;
    const EvalMod<word> &em = ctx->GetEvalMod();
    Msg in(slots), plain(slots), ideal(slots);
    for (int i = 0; i < slots; i++) {
      const double I = (double)((int)rng.UniformInt(0, 20) - 10);
      const double r = 0.01 * (2 * rng.UniformReal() - 1);
      in[i] = (I + r) / K;
      double y = em.Poly().PlainEvaluate(in[i].real() - 0.25 / K);
      double c = std::pow(0.5 / M_PI, 1.0 / (1 << bp.num_double_angle_));
      for (int d = 0; d < bp.num_double_angle_; d++) {
        c *= c;
        y = 2 * y * y - c;
      }
      plain[i] = y;
      ideal[i] = std::sin(2 * M_PI * r) / (2 * M_PI);
    }
    Ct ct, out;
    encrypt(ct, in, bp.GetEvalModStartLevel(), em.start_scale_);
    ctx->EvaluateMod(out, ct, ui.GetMultiplicationKey());
    const Msg got = decrypt(out);
    Report("EvalMod vs plain schedule", plain, got, 1e-6);
    Report("EvalMod vs sin(2 pi r) / (2 pi)", ideal, got, 1e-4);
  } catch (const std::exception &e) {
    std::printf("FAIL EvalMod: %s\n", e.what());
    g_fail++;
  }

  try {
    const auto &fft = ctx->GetEvalSpecialFFT(slots);
    for (int which = 0; which < 2; which++) {
      const auto &lt = which == 0 ? fft.CtSPhases().at(0) : fft.StCPhases().at(0);
      Msg m(slots);
      rng.SampleUniformComplex(m.data(), slots, -1.0, 1.0);
      const int level = lt.PtLevel();
      Ct ct, out;
      encrypt(ct, m, level, P.GetRescalePrimeProd(level));
      lt.Evaluate(*ctx, out, ct, keys);
      // a fresh encryption at a ~2^30 scale (bootparam_30) carries about
      // 2^10 times the relative noise of a ~2^40 one: below 2^35 the stage
      // is held to the reference's end-to-end tolerance instead
      const double scale = P.GetRescalePrimeProd(level);
      Report(which == 0 ? "CoeffToSlot phase 0 vs plain" : "SlotToCoeff phase 0 vs plain",
             lt.PlainEvaluate(m), decrypt(out), std::log2(scale) < 35 ? 1e-3 : 1e-4);
    }
  } catch (const std::exception &e) {
    std::printf("FAIL linear transform: %s\n", e.what());
    g_fail++;
  }

  try {
    // a hoisted transform with more baby steps (20) than one giant-step
    // accumulation pass takes (16) and with missing diagonals
    const int level = ctx->boot_param_.GetCtSStartLevel();
    PlainHoistMap map;
    Random r2(0x4348454444415207ull);
    for (int g : {0, 20, 40})
      for (int b = 0; b < 20; b++) {
        if ((b + g) % 7 == 3) continue;
        Msg d(slots);
        r2.SampleUniformComplex(d.data(), slots, -0.1, 0.1);
        map[g][b] = d;
      }
    HoistHandler<word> h(*ctx, map, level, P.GetRescalePrimeProd(level));
    EvkRequest rq;
    h.AddRequiredRotations(rq);
    ui.PrepareRotationKey(rq);
    Msg m(slots);
    rng.SampleUniformComplex(m.data(), slots, -1.0, 1.0);
    Ct ct, out;
    encrypt(ct, m, level, P.GetRescalePrimeProd(level));
    h.Evaluate(*ctx, out, ct, ui.GetEvkMap());
    Report("hoisted transform, 20 baby steps x 3 giant steps", h.PlainEvaluate(m), decrypt(out),
           1e-4);
  } catch (const std::exception &e) {
    std::printf("FAIL hoisted transform: %s\n", e.what());
    g_fail++;
  }

  try {
    Msg m(slots);
    rng.SampleUniformComplex(m.data(), slots, -1.0, 1.0);
    Ct ct, out;
    encrypt(ct, m, 0, P.GetScale(0));
    ctx->Boot(out, ct, keys);  // warm-up
    (void)hipDeviceSynchronize();
    double total = 0;
    for (int r = 0; r < runs; r++) {
      encrypt(ct, m, 0, P.GetScale(0));
      (void)hipDeviceSynchronize();
      auto t = Clock::now();
      ctx->Boot(out, ct, keys);
      (void)hipDeviceSynchronize();
      total += Seconds(t);
    }
    std::printf("boot: %.2f ms per bootstrap (%d runs), output level %d\n",
                1e3 * total / runs, runs, P.NPToLevel(out.GetNP()));
    Report("Boot (full slots)", m, decrypt(out), 1e-3);
    ctx->Boot(out, ct, keys, true);  // warm-up of the min-KS schedule
    (void)hipDeviceSynchronize();
    auto t = Clock::now();
    ctx->Boot(out, ct, keys, true);
    (void)hipDeviceSynchronize();
    std::printf("boot (min-KS): %.2f ms per bootstrap\n", 1e3 * Seconds(t));
    Report("Boot min-KS (full slots)", m, decrypt(out), 1e-3);
    if (P.IsUsingSparseSecretEncapsulation()) {
      // the non-default dense -> sparse switch with exact CRT lifts
      ctx->SetSseExactLift(true);
      ctx->Boot(out, ct, keys);
      ctx->SetSseExactLift(false);
      Report("Boot, exact-lift dense -> sparse switch (full slots)", m, decrypt(out), 1e-3);
    }
  } catch (const std::exception &e) {
    std::printf("FAIL Boot: %s\n", e.what());
    g_fail++;
  }
  try {
    const int batch = 4;
    std::vector<Ct> in(batch), out(batch), eager(batch);
    std::vector<Msg> msgs(batch, Msg(slots));
    for (int i = 0; i < batch; i++) {
      rng.SampleUniformComplex(msgs[i].data(), slots, -1.0, 1.0);
      encrypt(in[i], msgs[i], 0, P.GetScale(0));
      ctx->Boot(eager[i], in[i], keys);
    }
    hipStream_t st;
    HipCheck(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    {
      std::vector<const Ct *> ip;
      std::vector<Ct *> op;
      for (int i = 0; i < batch; i++) {
        ip.push_back(&in[i]);
        op.push_back(&out[i]);
      }
      HipCheck(hipDeviceSynchronize(), "sync");
      BootGraph<word> graph(*ctx, ip, op, keys, st);
      std::printf("graph: %d bootstraps, %zu nodes\n", batch, graph.Nodes());
      for (int rep = 0; rep < 3; rep++) {
        Ct scratch;
        ctx->Boot(scratch, in[0], keys);  // eager work between replays
        HipCheck(hipDeviceSynchronize(), "sync");
        graph.Launch();
        HipCheck(hipStreamSynchronize(st), "sync");
        bool same = true;
        for (int i = 0; i < batch; i++)
          same = same && out[i].bx_.CopyToHost() == eager[i].bx_.CopyToHost() &&
                 out[i].ax_.CopyToHost() == eager[i].ax_.CopyToHost();
        std::printf("%s graph replay %d equals the eager bootstraps\n", same ? "PASS" : "FAIL",
                    rep);
        if (!same) g_fail++;
      }
      auto t = Clock::now();
      graph.Launch();
      HipCheck(hipStreamSynchronize(st), "sync");
      std::printf("graph: %.2f ms per bootstrap (%d in one graph)\n",
                  1e3 * Seconds(t) / batch, batch);
      for (int i = 0; i < batch; i++)
        Report("graph bootstrap " + std::to_string(i), msgs[i], decrypt(out[i]), 1e-3);
    }
    HipCheck(hipStreamDestroy(st), "hipStreamDestroy");
  } catch (const std::exception &e) {
    std::printf("FAIL BootGraph: %s\n", e.what());
    g_fail++;
  }
  try {
    struct Case {
      int slots;
      BootVariant variant;
      const char *name;
    };
    for (const Case &c : {Case{slots / 2, BootVariant::kNormal, "Boot (2^14 slots)"},
                          Case{slots / 4, BootVariant::kImaginaryRemoving,
                               "Boot imaginary-removing (2^13 slots)"},
                          Case{slots / 8, BootVariant::kMergeTwoReal,
                               "Boot merge-two-real (2^12 slots), x + conj(x)"}}) {
      ctx->PrepareEvalSpecialFFT(c.slots, c.variant);
      EvkRequest r2;
      ctx->AddRequiredRotations(r2, c.slots);
      ui.PrepareRotationKey(r2);
      Msg m(c.slots);
      rng.SampleUniformComplex(m.data(), c.slots, -1.0, 1.0);
      Ct ct, out;
      encrypt(ct, m, 0, P.GetScale(0));
      ctx->Boot(out, ct, ui.GetEvkMap());
      Msg want = m, got = decrypt(out);
      if (c.variant != BootVariant::kNormal)
        for (auto &v : want) v = Complex(v.real(), 0);
      if (c.variant == BootVariant::kMergeTwoReal)
        for (auto &v : got) v = Complex(2 * v.real(), 0);  // x + conj(x)
      Report(c.name, want, got, 1e-3);
    }
  } catch (const std::exception &e) {
    std::printf("FAIL sparse-slot bootstrap: %s\n", e.what());
    g_fail++;
  }
  std::printf("OK: %d failure(s)\n", g_fail);
  return g_fail ? 1 : 0;
}
