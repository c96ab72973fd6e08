// Batched-plan benchmark over the C ABI alone (no torch in the process, so
// hardware-counter collectors can profile it): the 64-polynomial x 32-limb
// contractions at N = 2^16 on a reference-schema parameter file with at
// least 24 main and 8 auxiliary primes.
//
//   plan_bench <params.json> ntt   <valu|int8|auto> [reps] [blocks] [chunk]
//   plan_bench <params.json> bconv <fp64|int8|valu|auto> [reps]
//
// ntt: forward NTT of 64 polynomials x (24 main + 8 auxiliary) limbs, one
// plan (ck_plan_ntt; blocks / chunk: CK_OPT_NTT8_BLOCKS / CK_OPT_NTT8_CHUNK
// of the int8 form). bconv: BaseConv of 64 polynomials from the 8 auxiliary
// primes to all 32 (ck_plan_bconv). Inputs are seeded uniform residues
// (canonical for ntt, centred for bconv) uploaded once; the plan runs once
// untimed, then `reps` times between two HIP events on one stream. Prints
// one JSON object: ms per run, the plan's form and, when autotuned, every
// candidate's time.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cheddar/ParameterJson.h"
#include "cheddar_amd.h"

using namespace cheddar;

namespace {

void die(const char *what) {
  std::fprintf(stderr, "plan_bench: %s: %s\n", what, ck_last_error());
  std::exit(1);
}
void hip_ok(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "plan_bench: %s: %s\n", what, hipGetErrorString(e));
    std::exit(1);
  }
}

uint64_t g_state = 0x4348454444415203ull;
uint64_t next_u64() {  // SplitMix64
  uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int form_code(const std::string &w, const std::string &f) {
  if (f == "auto") return -1;
  if (w == "ntt") return f == "int8" ? 1 : 0;
  return f == "int8" ? 1 : f == "valu" ? 2 : 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr,
                 "usage: %s <params.json> ntt <valu|int8|auto> [reps] [blocks] [chunk]\n"
                 "       %s <params.json> bconv <fp64|int8|valu|auto> [reps]\n",
                 argv[0], argv[0]);
    return 2;
  }
  const std::string work = argv[2], form = argv[3];
  const int reps = argc > 4 ? std::atoi(argv[4]) : 20;
  const int blocks = argc > 5 ? std::atoi(argv[5]) : 16;
  const int chunk = argc > 6 ? std::atoi(argv[6]) : 0;
  auto prm = ParameterFromJsonFile(argv[1]);
  const Parameter<uint32_t> &P = *prm;
  std::vector<int> cfg;
  for (auto &c : P.level_config_) {
    cfg.push_back(c.first);
    cfg.push_back(c.second);
  }
  ck_context *ctx = nullptr;
  if (ck_context_create(P.log_degree_, P.main_primes_.data(), (int)P.main_primes_.size(),
                        P.ter_primes_.data(), (int)P.ter_primes_.size(), P.aux_primes_.data(),
                        (int)P.aux_primes_.size(), cfg.data(), (int)cfg.size() / 2, &ctx))
    die("ck_context_create");
  const int n = P.degree_, npoly = 64, nm = 24, na = 8;
  if ((int)P.main_primes_.size() < nm || (int)P.aux_primes_.size() < na ||
      !P.ter_primes_.empty()) {
    std::fprintf(stderr, "plan_bench: needs >= 24 main, >= 8 auxiliary, no terminal primes\n");
    return 2;
  }
  // global prime order [main | aux]: the 32 primes of the 24 + 8 limbs
  std::vector<uint32_t> primes(P.main_primes_.begin(), P.main_primes_.begin() + nm);
  primes.insert(primes.end(), P.aux_primes_.begin(), P.aux_primes_.begin() + na);
  const int aux0 = (int)P.main_primes_.size();
  const bool is_ntt = work == "ntt";
  if (!is_ntt && work != "bconv") {
    std::fprintf(stderr, "plan_bench: unknown workload %s\n", work.c_str());
    return 2;
  }
  const int in_limbs = is_ntt ? nm + na : na, out_limbs = nm + na;
  std::vector<uint32_t *> src(npoly), dst(npoly);
  std::vector<uint32_t> host((size_t)in_limbs * n);
  for (int k = 0; k < npoly; k++) {
    for (int y = 0; y < in_limbs; y++) {
      const uint32_t q = is_ntt ? primes[y] : primes[nm + y];
      for (int i = 0; i < n; i++) {
        const uint32_t r = (uint32_t)(next_u64() % q);
        host[(size_t)y * n + i] = is_ntt ? r : (uint32_t)((int32_t)r - (int32_t)(q / 2));
      }
    }
    hip_ok(hipMalloc((void **)&src[k], host.size() * 4), "hipMalloc");
    hip_ok(hipMalloc((void **)&dst[k], (size_t)out_limbs * n * 4), "hipMalloc");
    hip_ok(hipMemcpy(src[k], host.data(), host.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
  }
  const int code = form_code(work, form);
  ck_plan *plan = nullptr;
  if (is_ntt) {
    if (ck_context_set_option(ctx, CK_OPT_PLAN_NTT_FORM, code) ||
        ck_context_set_option(ctx, CK_OPT_NTT8_BLOCKS, blocks) ||
        ck_context_set_option(ctx, CK_OPT_NTT8_CHUNK, chunk))
      die("options");
    if (ck_plan_ntt(ctx, npoly, nm, 0, na, 0, 0, (const uint32_t *const *)src.data(),
                    dst.data(), &plan))
      die("ck_plan_ntt");
  } else {
    if (ck_context_set_option(ctx, CK_OPT_PLAN_BCONV_FORM, code)) die("options");
    std::vector<int> sp, dp;
    for (int i = 0; i < na; i++) sp.push_back(aux0 + i);
    for (int i = 0; i < nm; i++) dp.push_back(i);
    for (int i = 0; i < na; i++) dp.push_back(aux0 + i);
    if (ck_plan_bconv(ctx, sp.data(), na, dp.data(), out_limbs, npoly,
                      (const int32_t *const *)src.data(), dst.data(), &plan))
      die("ck_plan_bconv");
  }
  hipStream_t st;
  hipEvent_t e0, e1;
  hip_ok(hipStreamCreate(&st), "hipStreamCreate");
  hip_ok(hipEventCreate(&e0), "hipEventCreate");
  hip_ok(hipEventCreate(&e1), "hipEventCreate");
  if (ck_plan_run(plan, st)) die("ck_plan_run");
  hip_ok(hipStreamSynchronize(st), "sync");
  hip_ok(hipEventRecord(e0, st), "record");
  for (int r = 0; r < reps; r++)
    if (ck_plan_run(plan, st)) die("ck_plan_run");
  hip_ok(hipEventRecord(e1, st), "record");
  hip_ok(hipEventSynchronize(e1), "sync");
  float ms = 0;
  hip_ok(hipEventElapsedTime(&ms, e0, e1), "elapsed");
  long long info[8];
  ck_plan_info(plan, info);
  std::printf("{\"workload\": \"%s\", \"form\": \"%s\", \"plan_form\": %lld, \"reps\": %d, "
              "\"ms\": %.5f, \"launches\": %lld",
              work.c_str(), form.c_str(), info[7], reps, ms / reps, info[2]);
  const int nt = ck_plan_tuning(plan, 0, nullptr, nullptr);
  if (nt > 0) {
    std::vector<int> f(3 * nt);
    std::vector<double> us(nt);
    ck_plan_tuning(plan, nt, f.data(), us.data());
    std::printf(", \"tuning\": [");
    for (int i = 0; i < nt; i++)
      std::printf("%s[%d, %d, %d, %.2f]", i ? ", " : "", f[3 * i], f[3 * i + 1], f[3 * i + 2],
                  us[i]);
    std::printf("]");
  }
  std::printf("}\n");
  ck_plan_destroy(plan);
  for (int k = 0; k < npoly; k++) {
    (void)hipFree(src[k]);
    (void)hipFree(dst[k]);
  }
  ck_context_destroy(ctx);
  return 0;
}
