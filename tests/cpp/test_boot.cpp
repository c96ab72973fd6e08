// Bootstrapping through the C++ host facade on the GPU (the reference's
// unittest/Bootstrapping.cpp: encrypt a random message at level 0,
// bootstrap, decrypt and compare within 1e-3), with stage checks first:
// - ModUpToMax word for word against the oracle (with encapsulation: the
//   dense -> sparse switch at level -1 and the sparse -> dense switch);
// - EvalMod on values (I + r) / K against the plain evaluation of the same
//   schedule and against sin(2 pi r) / (2 pi);
// - the first CoeffToSlot and SlotToCoeff phases against the plain
//   baby-step / giant-step product of their plaintext maps;
// - the full bootstrap: precision bits = -log2(max error) and the mean
//   error, and the wall time per bootstrap (after one warm-up); against an
//   FP64 plaintext simulation of the same schedule from the same raised
//   input (the precision ceiling; within 0.5 bits where the schedule's
//   approximation error, not the homomorphic noise, sets the precision);
//   the same bootstrap with the exact-lift dense -> sparse switch;
// - four bootstraps captured into one HIP graph: three replays, with eager
//   work and synchronisation between them, each output equal word for word
//   to the eager bootstrap of the same input (the eager bootstraps run
//   their pairs of independent steps on two streams, a graph captures them
//   in order: the comparison also checks that the pairs change nothing);
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

// Error bits of got against want: -log2(max |error|), -log2(mean |error|)
std::pair<double, double> ErrorBits(const Msg &want, const Msg &got) {
  double worst = 0, sum = 0;
  for (size_t i = 0; i < want.size(); i++) {
    const double e = std::abs(want[i] - got[i]);
    worst = std::max(worst, e);
    sum += e;
  }
  return {-std::log2(worst), -std::log2(sum / want.size())};
}

// FP64 plaintext simulation of Boot at full slots (BootContext.cpp:299-365):
// the same encoding (truncation at scale base_scale), the level-zero
// scale-up, the same CoeffToSlot / SlotToCoeff phases (their plain maps
// carry the same constants) with the same scale reinterpretations, and the
// same EvalMod schedule (series, then double-angle steps), with no
// encryption noise and no rounding of rescales. The modulus raise is the
// exact integer lift of a level-0 encryption under a secret of the sparse
// Hamming weight: t = b~ + a~ s with b~, a~ centred modulo Q0, so
// t = 2^s c + Q0 I, the input EvalMod sees in the real bootstrap. With
// `raised` (the slots of a decrypted ScaleUpAndRaise output) the
// simulation starts from that t instead, so that it sees the same integers
// I as the bootstrap it is compared with.
Msg SimulateBoot(const BootContext<word> &ctx, const Msg &m, uint64_t seed,
                 const Msg *raised = nullptr) {
  const Parameter<word> &P = ctx.param_;
  const BootParameter &bp = ctx.boot_param_;
  const int n = P.degree_, half = n / 2, M = 2 * n;
  std::vector<Complex> tw(M);
  for (int i = 0; i < M; i++) tw[i] = ctx.encoder_.GetTwiddleFactor(i);
  std::vector<int> gal(half + 1);
  for (int j = 0; j <= half; j++) gal[j] = P.GetGaloisFactor(j);
  // 1. encode at level 0 (truncation, Encode.cpp:129-182), scale up
  const double d0 = P.GetScale(0);
  std::vector<Complex> z(m);
  SpecialIFFT(z, tw, gal, M);
  std::vector<long double> t(n);
  const long double up = std::ldexp(1.0L, ctx.GetLogScaleUp());
  for (int i = 0; i < half; i++) {
    t[i] = std::trunc(z[i].real() * d0) * up;
    t[i + half] = std::trunc(z[i].imag() * d0) * up;
  }
  // 2. the lift: b~ = [2^s c - a~ s]_Q0 centred, t = b~ + a~ s
  long double q0 = 1;
  for (word q : P.GetPrimeVector(P.LevelToNP(-1))) q0 *= q;
  Random r(seed);
  std::vector<long double> a(n), as(n, 0);
  for (int i = 0; i < n; i++) a[i] = std::floor(((long double)r.UniformReal() - 0.5L) * q0);
  std::vector<char> used(n, 0);
  for (int k = 0; k < P.GetSparseHammingWeight(); k++) {
    int pos;
    do pos = (int)r.UniformInt(0, n - 1);
    while (used[pos]);
    used[pos] = 1;
    const long double sgn = r.UniformInt(0, 1) ? 1 : -1;
    for (int i = 0; i < n; i++) {  // a~ (+-X^pos), negacyclic
      const int j = i + pos;
      if (j < n) as[j] += sgn * a[i];
      else as[j - n] -= sgn * a[i];
    }
  }
  for (int i = 0; i < n; i++) {
    long double b = std::fmod(t[i] - as[i], q0);
    if (b > q0 / 2) b -= q0;
    if (b <= -q0 / 2) b += q0;
    t[i] = b + as[i];
  }
  // 3. the slots at the tracked scale (base_scale: the scale-up constant
  // has scale 1), CoeffToSlot, EvalMod's start scale
  std::vector<Complex> u(half);
  for (int i = 0; i < half; i++)
    u[i] = Complex((double)(t[i] / d0), (double)(t[i + half] / d0));
  SpecialFFT(u, tw, gal, M);
  if (raised) u = *raised;
  const auto &fft = ctx.GetEvalSpecialFFT(half);
  for (const auto &ph : fft.CtSPhases()) u = ph.PlainEvaluate(u);
  const EvalMod<word> &em = ctx.GetEvalMod();
  const double in_f = d0 / em.start_scale_;
  const int K = (1 << bp.num_double_angle_) * bp.initial_K_;
  auto eval_mod = [&](double x) {
    double y = em.Poly().PlainEvaluate(x - 0.25 / K);
    double c = std::pow(0.5 / M_PI, 1.0 / (1 << bp.num_double_angle_));
    for (int d = 0; d < bp.num_double_angle_; d++) {
      c *= c;
      y = 2 * y * y - c;
    }
    return y;
  };
  // 4. real part w + conj(w), imaginary part (conj(w) - w) i, recombined
  for (auto &w : u) {
    const Complex v = w * in_f;
    w = Complex(eval_mod(2 * v.real()), eval_mod(2 * v.imag()));
  }
  // 5. SlotToCoeff, the end scale reinterpreted as the end level's
  for (const auto &ph : fft.StCPhases()) u = ph.PlainEvaluate(u);
  const double out_f = em.end_scale_ / P.GetScale(bp.GetEndLevel());
  for (auto &w : u) w *= out_f;
  return u;
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
    const Msg got = decrypt(out);
    Report("Boot (full slots)", m, got, 1e-3);
    // precision ceiling: the FP64 plaintext simulation of the same schedule,
    // from an independent lift and from this bootstrap's own raised input
    Ct raised;
    ctx->ScaleUpAndRaise(raised, ct, keys);
    raised.SetNumSlots(slots);
    const Msg raised_slots = decrypt(raised);
    const Msg sim = SimulateBoot(*ctx, m, 0x4348454444415208ull, &raised_slots);
    const Msg sim_free = SimulateBoot(*ctx, m, 0x4348454444415208ull);
    const auto [gpu_max, gpu_mean] = ErrorBits(m, got);
    const auto [sim_max, sim_mean] = ErrorBits(m, sim);
    const auto [free_max, free_mean] = ErrorBits(m, sim_free);
    const auto [noise_max, noise_mean] = ErrorBits(sim, got);
    std::printf("FP64 simulation from an independent lift: %.2f bits max (%.2f mean)\n",
                free_max, free_mean);
    std::printf("homomorphic noise (GPU - simulation): %.2f bits max (%.2f mean)\n",
                noise_max, noise_mean);
    // The ceiling binds where the schedule's approximation error exceeds the
    // homomorphic noise: a 2^40 scale with a sparse main secret. Otherwise
    // (2^30 / 2^35 scales, or the dense secret of encapsulation, whose
    // rescale rounding noise is sqrt(h) times larger) the noise printed
    // above sets the precision and the gap is reported only.
    const bool binding =
        !P.IsUsingSparseSecretEncapsulation() && std::log2(P.GetScale(0)) >= 39.5;
    const bool within = gpu_max >= sim_max - 0.5;
    std::printf("%s Boot precision vs the FP64 plaintext simulation of the same input: GPU "
                "%.2f bits max (%.2f mean), simulation %.2f bits max (%.2f mean), gap %.2f "
                "bits%s\n",
                !binding ? "INFO" : within ? "PASS" : "FAIL", gpu_max, gpu_mean, sim_max,
                sim_mean, sim_max - gpu_max, binding ? " (bound 0.5)" : "");
    std::fflush(stdout);
    if (binding && !within) g_fail++;
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
    {
      // two lanes: a second context of the same parameters, bootstraps 1
      // and 3 on its own stream inside the same graph
      auto lane = BootContext<word>::Create(P, BootParameter(P.max_level_, ncts, nstc));
      lane->PrepareEvalMod();
      lane->PrepareEvalSpecialFFT(slots);
      Ct warm;
      lane->Boot(warm, in[0], keys);
      std::vector<Ct> out2(batch);
      std::vector<const Ct *> ip;
      std::vector<Ct *> op;
      for (int i = 0; i < batch; i++) {
        ip.push_back(&in[i]);
        op.push_back(&out2[i]);
      }
      HipCheck(hipDeviceSynchronize(), "sync");
      BootGraph<word> graph({ctx.get(), lane.get()}, ip, op, keys, st);
      for (int rep = 0; rep < 2; rep++) {
        graph.Launch();
        HipCheck(hipStreamSynchronize(st), "sync");
        bool same = true;
        for (int i = 0; i < batch; i++)
          same = same && out2[i].bx_.CopyToHost() == eager[i].bx_.CopyToHost() &&
                 out2[i].ax_.CopyToHost() == eager[i].ax_.CopyToHost();
        std::printf("%s two-lane graph replay %d equals the eager bootstraps (%zu lanes)\n",
                    same ? "PASS" : "FAIL", rep, graph.Lanes());
        if (!same) g_fail++;
      }
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
