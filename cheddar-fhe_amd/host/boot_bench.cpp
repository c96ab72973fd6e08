// Bootstrapping benchmark over the C++ host facade: builds a BootContext
// for a reference-schema parameter file (full slot count), generates its
// keys, then times `runs` bootstraps of a ciphertext encrypting a message
// uniform in the unit box (the reference test's distribution, after one
// warm-up), checks the result and prints one JSON object: ms per
// bootstrap, precision bits (-log2 of the max and the mean error), set-up
// time, the same for 8 ciphertexts with messages uniform in the unit disk
// bootstrapped by one captured HIP graph over `lanes` BootContexts (ms per
// bootstrap over the graph's launch; precision over all 8; the outputs
// compared word for word with eager bootstraps of the same inputs), and,
// with profiling on for one extra bootstrap, the summed GPU time per kernel.
//
//   boot_bench <bootparam.json> [runs] [seed] [sse|nosse] [bconv] [lanes]
//
// sse (default): the file's dense Hamming weight with sparse-secret
// encapsulation; nosse: the main secret gets the sparse weight. bconv: the
// CK_OPT_BCONV_MFMA value (-1 default, 0 fused, 1 FP64, 2 int8 matrix
// cores). lanes (default 4): contexts of the graph, bootstrap i on lane
// i mod lanes, each lane a stream of its own.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "cheddar/ParameterJson.h"
#include "cheddar/UserInterface.h"
#include "cheddar/extension/BootContext.h"
#include "cheddar/extension/BootGraph.h"

using namespace cheddar;
using word = uint32_t;
using Clock = std::chrono::steady_clock;

static double Since(Clock::time_point t) {
  return std::chrono::duration<double>(Clock::now() - t).count();
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <bootparam.json> [runs] [seed]\n", argv[0]);
    return 2;
  }
  const int runs = argc > 2 ? std::atoi(argv[2]) : 3;
  const uint64_t seed = argc > 3 ? std::strtoull(argv[3], nullptr, 0) : 0x4348454444415204ull;
  const bool sse = !(argc > 4 && std::string(argv[4]) == "nosse");
  try {
    std::ifstream f(argv[1]);
    if (!f) throw Error(std::string("cannot open ") + argv[1]);
    std::stringstream ss;
    ss << f.rdbuf();
    auto param = ParameterFromJsonText(ss.str());
    int ncts = 0, nstc = 0;
    if (!BootLevelsFromJsonText(ss.str(), &ncts, &nstc))
      throw Error("the parameter file does not enable bootstrapping");
    if (!sse) param->SetDenseHammingWeight(param->GetSparseHammingWeight());
    const Parameter<word> &P = *param;
    const int slots = P.degree_ / 2;

    const int lanes = argc > 6 ? std::max(1, std::atoi(argv[6])) : 4;
    auto t0 = Clock::now();
    auto ctx = BootContext<word>::Create(P, BootParameter(P.max_level_, ncts, nstc));
    if (argc > 5) ctx->SetOption(CK_OPT_BCONV_MFMA, std::atoi(argv[5]));
    UserInterface<word> ui(ctx, seed);
    ctx->PrepareEvalMod();
    ctx->PrepareEvalSpecialFFT(slots);
    EvkRequest req;
    ctx->AddRequiredRotations(req, slots);
    ctx->AddRequiredRotations(req, slots, true);  // the min-KS strides
    ui.PrepareRotationKey(req);
    HipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const double setup = Since(t0);

    Random rng(seed + 1);
    std::vector<Complex> msg(slots);
    rng.SampleUniformComplex(msg.data(), slots, -1.0, 1.0);
    Plaintext<word> pt;
    ctx->encoder_.Encode(pt, 0, P.GetScale(0), msg);
    Ciphertext<word> ct, out;
    ui.Encrypt(ct, pt);
    ctx->Boot(out, ct, ui.GetEvkMap());  // warm-up
    HipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    double total = 0;
    for (int r = 0; r < runs; r++) {
      auto t = Clock::now();
      ctx->Boot(out, ct, ui.GetEvkMap());
      HipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
      total += Since(t);
    }
    Plaintext<word> dec;
    ui.Decrypt(dec, out);
    std::vector<Complex> got;
    ctx->encoder_.Decode(got, dec);
    double worst = 0, sum = 0;
    for (int i = 0; i < slots; i++) {
      const double e = std::abs(got[i] - msg[i]);
      if (!(e <= worst)) worst = e;
      sum += e;
    }

    // the minimum-key-switching schedule
    ctx->Boot(out, ct, ui.GetEvkMap(), true);  // warm-up
    HipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    double minks_total = 0;
    for (int r = 0; r < runs; r++) {
      auto t = Clock::now();
      ctx->Boot(out, ct, ui.GetEvkMap(), true);
      HipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
      minks_total += Since(t);
    }
    double minks_worst = 0;
    {
      Plaintext<word> d;
      ui.Decrypt(d, out);
      std::vector<Complex> g;
      ctx->encoder_.Decode(g, d);
      for (int j = 0; j < slots; j++) {
        const double e = std::abs(g[j] - msg[j]);
        if (!(e <= minks_worst)) minks_worst = e;
      }
    }

    // `batch` bootstraps in one captured graph over `lanes` contexts (the
    // first is ctx; the others prepared alike, each with its own workspace)
    const int batch = 8;
    auto tl = Clock::now();
    std::vector<std::shared_ptr<BootContext<word>>> extra;
    std::vector<BootContext<word> *> lane_ctx{ctx.get()};
    for (int j = 1; j < lanes; j++) {
      extra.push_back(BootContext<word>::Create(P, BootParameter(P.max_level_, ncts, nstc)));
      if (argc > 5) extra.back()->SetOption(CK_OPT_BCONV_MFMA, std::atoi(argv[5]));
      extra.back()->PrepareEvalMod();
      extra.back()->PrepareEvalSpecialFFT(slots);
      Ciphertext<word> warm;
      extra.back()->Boot(warm, ct, ui.GetEvkMap());  // warms the pool for the capture
      lane_ctx.push_back(extra.back().get());
    }
    HipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const double lanes_setup = Since(tl);
    std::vector<Ciphertext<word>> gin(batch), gout(batch), geager(batch);
    std::vector<std::vector<Complex>> gmsg(batch, std::vector<Complex>(slots));
    for (int i = 0; i < batch; i++) {
      // uniform in the unit disk: radius sqrt(u), angle 2 pi v
      for (auto &z : gmsg[i]) z = std::polar(std::sqrt(rng.UniformReal()), 2 * M_PI * rng.UniformReal());
      Plaintext<word> gp;
      ctx->encoder_.Encode(gp, 0, P.GetScale(0), gmsg[i]);
      ui.Encrypt(gin[i], gp);
      ctx->Boot(geager[i], gin[i], ui.GetEvkMap());
    }
    hipStream_t st;
    HipCheck(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    double graph_ms = 0;
    size_t graph_nodes = 0;
    double graph_worst = 0, graph_sum = 0;
    bool graph_same = true;
    {
      std::vector<const Ciphertext<word> *> ip;
      std::vector<Ciphertext<word> *> op;
      for (int i = 0; i < batch; i++) {
        ip.push_back(&gin[i]);
        op.push_back(&gout[i]);
      }
      HipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
      BootGraph<word> graph(lane_ctx, ip, op, ui.GetEvkMap(), st);
      graph_nodes = graph.Nodes();
      graph.Launch();  // warm-up
      HipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
      auto t = Clock::now();
      for (int r = 0; r < runs; r++) graph.Launch();
      HipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
      graph_ms = 1e3 * Since(t) / (runs * batch);
      for (int i = 0; i < batch; i++)
        graph_same = graph_same && gout[i].bx_.CopyToHost() == geager[i].bx_.CopyToHost() &&
                     gout[i].ax_.CopyToHost() == geager[i].ax_.CopyToHost();
      for (int i = 0; i < batch; i++) {
        Plaintext<word> d;
        ui.Decrypt(d, gout[i]);
        std::vector<Complex> g;
        ctx->encoder_.Decode(g, d);
        for (int j = 0; j < slots; j++) {
          const double e = std::abs(g[j] - gmsg[i][j]);
          if (!(e <= graph_worst)) graph_worst = e;
          graph_sum += e;
        }
      }
    }
    HipCheck(hipStreamDestroy(st), "hipStreamDestroy");

    // one more bootstrap with per-kernel timing
    // (its EvalMod pair in order, so every launch is on the profiled context)
    ctx->SetPairOverlap(false);
    Context<word>::Check(ck_profile_enable(ctx->Handle(), 1));
    ctx->Boot(out, ct, ui.GetEvkMap());
    constexpr int kMax = 1024;
    std::vector<char> names(64 * kMax);
    std::vector<double> ms(kMax);
    std::vector<long long> calls(kMax);
    const int n = ck_profile_report(ctx->Handle(), kMax, names.data(), ms.data(), calls.data());
    Context<word>::Check(ck_profile_enable(ctx->Handle(), 0));
    ctx->SetPairOverlap(true);

    std::printf("{\"ms_per_bootstrap\": %.3f, \"runs\": %d, \"precision_bits_max\": %.2f, "
                "\"precision_bits_mean\": %.2f, \"max_error\": %.3e, \"setup_s\": %.1f, "
                "\"rotation_keys\": %zu, \"slots\": %d, \"end_level\": %d, "
                "\"dense_hamming_weight\": %d, \"sparse_hamming_weight\": %d, "
                "\"sparse_secret_encapsulation\": %s, \"graph_batch\": %d, "
                "\"graph_ms_per_bootstrap\": %.3f, \"graph_nodes\": %zu, "
                "\"graph_lanes\": %d, \"graph_lanes_setup_s\": %.1f, "
                "\"graph_equal_to_eager\": %s, "
                "\"graph_precision_bits_max\": %.2f, \"graph_precision_bits_mean\": %.2f, "
                "\"graph_message\": \"uniform in the unit disk\", \"minks_ms_per_bootstrap\": %.3f, "
                "\"minks_precision_bits_max\": %.2f, \"kernels\": {",
                1e3 * total / runs, runs, -std::log2(worst), -std::log2(sum / slots), worst,
                setup, req.size(), slots, P.NPToLevel(out.GetNP()),
                P.GetDenseHammingWeight(), P.GetSparseHammingWeight(),
                P.IsUsingSparseSecretEncapsulation() ? "true" : "false", batch, graph_ms,
                graph_nodes, lanes, lanes_setup, graph_same ? "true" : "false",
                -std::log2(graph_worst), -std::log2(graph_sum / (batch * slots)),
                1e3 * minks_total / runs,
                -std::log2(minks_worst));
    for (int i = 0; i < n; i++)
      std::printf("%s\"%s\": [%.4f, %lld]", i ? ", " : "", &names[64 * i], ms[i], calls[i]);
    std::printf("}}\n");
    return worst < 1e-3 && graph_worst < 1e-3 && minks_worst < 1e-3 && graph_same ? 0 : 1;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "boot_bench: %s\n", e.what());
    return 1;
  }
}
