// C++ façade test: the reference's evaluator calls (Context<uint32_t>::
// HMult, Relinearize, Rescale, HRot, HConj, MultKeyNoModDown, Add/Sub/Neg,
// Mult + RelinearizeRescale; unittest/BasicTest.cpp:327-419 exercises the
// same calls) made through cheddar/Context.h on the GPU, every output
// compared word for word with the CPU oracle (oracle/ckks_oracle.h) on the
// same seeded inputs. Parameters come from a reference-schema JSON file.
//
//   test_context <params.json> [level]
//
// Prints one line per case and exits non-zero on the first mismatch.
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <vector>

#include "cheddar/Context.h"
#include "cheddar/ParameterJson.h"
#include "ckks_oracle.h"

using namespace cheddar;
using word = uint32_t;
using Ct = Ciphertext<word>;
using Evk = EvaluationKey<word>;
using Pt = Plaintext<word>;
using Const = Constant<word>;
using Host = std::vector<word>;

namespace {

int g_fail = 0;

void Report(const std::string &name, bool ok, const std::string &detail = "") {
  std::printf("%s %s%s%s\n", ok ? "PASS" : "FAIL", name.c_str(),
              detail.empty() ? "" : ": ", detail.c_str());
  std::fflush(stdout);
  if (!ok) g_fail++;
}

bool Same(const Host &got, const Host &want, std::string &detail) {
  if (got.size() != want.size()) {
    detail = "size " + std::to_string(got.size()) + " vs " + std::to_string(want.size());
    return false;
  }
  for (size_t i = 0; i < got.size(); i++)
    if (got[i] != want[i]) {
      detail = "first mismatch at word " + std::to_string(i);
      return false;
    }
  return true;
}

void Expect(const std::string &name, const DeviceVector<word> &got, const Host &want) {
  (void)hipDeviceSynchronize();
  std::string d;
  const bool ok = Same(got.CopyToHost(), want, d);
  Report(name, ok, d);
}

struct Harness {
  const Parameter<word> &param;
  or_param *op = nullptr;
  int n = 0;

  explicit Harness(const Parameter<word> &p) : param(p) {
    std::vector<int> cfg;
    for (auto &c : p.level_config_) {
      cfg.push_back(c.first);
      cfg.push_back(c.second);
    }
    op = or_param_create(p.log_degree_, p.main_primes_.data(), (int)p.main_primes_.size(),
                         p.ter_primes_.data(), (int)p.ter_primes_.size(),
                         p.aux_primes_.data(), (int)p.aux_primes_.size(), cfg.data(),
                         (int)p.level_config_.size());
    if (!op) throw Error("oracle rejected the parameters");
    n = p.degree_;
  }
  ~Harness() { or_param_destroy(op); }

  Host Uniform(const NPInfo &np, uint64_t seed) const {
    Host h((size_t)np.GetNumTotal() * n);
    or_fill_uniform(op, np.num_main_, np.num_ter_, np.num_aux_, seed, h.data());
    return h;
  }

  void Upload(Ct &ct, const NPInfo &np, uint64_t seed, Host &b, Host &a) const {
    ct.ModifyNP(np);
    b = Uniform(np, seed);
    a = Uniform(np, seed + 1);
    ct.bx_.CopyFromHost(b);
    ct.ax_.CopyFromHost(a);
  }

  // a uniform full-size key: dnum blocks of (L + alpha) limbs per part
  void MakeKey(Evk &key, uint64_t seed, Host &kb, Host &ka) const {
    const int dnum = param.dnum_;
    const NPInfo full(param.GetMaxNumMain(), param.GetMaxNumTer(), param.alpha_);
    key = Evk(full, dnum);
    const size_t blk = (size_t)full.GetNumTotal() * n;
    kb.assign(blk * dnum, 0);
    ka.assign(blk * dnum, 0);
    for (int g = 0; g < dnum; g++) {
      Host b = Uniform(full, seed + 2 * g), a = Uniform(full, seed + 2 * g + 1);
      std::copy(b.begin(), b.end(), kb.begin() + g * blk);
      std::copy(a.begin(), a.end(), ka.begin() + g * blk);
      key.bx_[g].CopyFromHost(b);
      key.ax_[g].CopyFromHost(a);
    }
  }

  std::vector<int> Primes(const NPInfo &np) const {
    std::vector<int> v;
    for (int t = 0; t < np.GetNumTotal(); t++)
      v.push_back(or_np_prime(op, np.num_main_, np.num_ter_, t));
    return v;
  }
};

void Run(const std::string &name, const std::function<void()> &f) {
  try {
    f();
  } catch (const std::exception &e) {
    Report(name, false, std::string("exception: ") + e.what());
  }
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <params.json> [level]\n", argv[0]);
    return 2;
  }
  or_set_num_threads(8);
  auto param = ParameterFromJsonFile(argv[1]);
  const int level = argc > 2 ? std::atoi(argv[2]) : param->max_level_;
  auto ctx = Context<word>::Create(*param);
  Harness H(*param);
  const int n = param->degree_;
  const NPInfo np = param->LevelToNP(level);
  const uint64_t seed = 0x4348454444415200ull + 900;
  std::printf("params %s: N=%d L=%d alpha=%d dnum=%d, level %d (%d q limbs)\n", argv[1], n,
              param->L_, param->alpha_, param->dnum_, level, np.GetNumQ());

  Ct c1, c2;
  Host b1, a1, b2, a2;
  H.Upload(c1, np, seed, b1, a1);
  H.Upload(c2, np, seed + 2, b2, a2);
  c1.SetScale(param->GetScale(std::min(level, param->default_encryption_level_)));
  c2.SetScale(c1.GetScale());
  Evk mk, rk, ck;
  Host mkb, mka, rkb, rka, ckb, cka;
  H.MakeKey(mk, seed + 10, mkb, mka);
  H.MakeKey(rk, seed + 40, rkb, rka);
  H.MakeKey(ck, seed + 70, ckb, cka);
  const size_t qs = (size_t)np.GetNumQ() * n;
  const std::vector<int> pidx = H.Primes(np);

  Run("Add / Sub / Neg", [&] {
    Ct r;
    Host want(qs);
    ctx->Add(r, c1, c2);
    or_add(H.op, pidx.data(), np.GetNumQ(), b1.data(), b2.data(), want.data());
    Expect("Add bx", r.bx_, want);
    ctx->Sub(r, c1, c2);
    or_sub(H.op, pidx.data(), np.GetNumQ(), a1.data(), a2.data(), want.data());
    Expect("Sub ax", r.ax_, want);
    ctx->Neg(r, c1);
    or_neg(H.op, pidx.data(), np.GetNumQ(), b1.data(), want.data());
    Expect("Neg bx", r.bx_, want);
  });

  Run("HMult", [&] {
    Ct r;
    ctx->HMult(r, c1, c2, mk, true);
    const NPInfo lo = param->LevelToNP(level - 1);
    Host wb((size_t)lo.GetNumQ() * n), wa(wb.size());
    or_hmult(H.op, level, b1.data(), a1.data(), b2.data(), a2.data(), mkb.data(),
             mka.data(), 1, wb.data(), wa.data());
    Expect("HMult (rescale) bx", r.bx_, wb);
    Expect("HMult (rescale) ax", r.ax_, wa);
    const double want_scale = c1.GetScale() * c2.GetScale() /
                              param->GetRescalePrimeProd(level);
    Report("HMult scale and level",
           r.GetNP() == lo && std::fabs(r.GetScale() / want_scale - 1.0) < 1e-12);
    // in place: res aliases an operand
    Ct t;
    ctx->Copy(t, c1);
    ctx->HMult(t, t, c2, mk, true);
    Expect("HMult in place bx", t.bx_, wb);
  });

  Run("Mult + Relinearize / RelinearizeRescale", [&] {
    Ct m, r;
    ctx->Mult(m, c1, c2);
    Report("Mult has rx", m.HasRx());
    ctx->RelinearizeRescale(r, m, mk);
    const NPInfo lo = param->LevelToNP(level - 1);
    Host wb((size_t)lo.GetNumQ() * n), wa(wb.size());
    or_hmult(H.op, level, b1.data(), a1.data(), b2.data(), a2.data(), mkb.data(),
             mka.data(), 1, wb.data(), wa.data());
    Expect("Mult+RelinearizeRescale bx", r.bx_, wb);
    Expect("Mult+RelinearizeRescale ax", r.ax_, wa);
    ctx->Relinearize(r, m, mk);
    Host nb(qs), na(qs);
    or_hmult(H.op, level, b1.data(), a1.data(), b2.data(), a2.data(), mkb.data(),
             mka.data(), 0, nb.data(), na.data());
    Expect("Mult+Relinearize ax", r.ax_, na);
    bool threw = false;
    try {
      ctx->HMult(r, m, c2, mk);
    } catch (const Error &) {
      threw = true;
    }
    Report("HMult rejects an unrelinearized operand", threw);
  });

  Run("Rescale", [&] {
    Ct r;
    ctx->Rescale(r, c1);
    const NPInfo lo = param->LevelToNP(level - 1);
    Host want((size_t)lo.GetNumQ() * n);
    or_moddown(H.op, level, OR_RESCALE, b1.data(), want.data());
    Expect("Rescale bx", r.bx_, want);
  });

  Run("HRot / HConj / hoisted", [&] {
    Ct r;
    Host wb(qs), wa(qs);
    for (int rot : {1, 4321}) {
      ctx->HRot(r, c1, rk, rot);
      or_hrot(H.op, level, b1.data(), a1.data(), rkb.data(), rka.data(), rot, wb.data(),
              wa.data());
      Expect("HRot " + std::to_string(rot) + " bx", r.bx_, wb);
      Expect("HRot " + std::to_string(rot) + " ax", r.ax_, wa);
    }
    ctx->HConj(r, c1, ck);
    or_hconj(H.op, level, b1.data(), a1.data(), ckb.data(), cka.data(), wb.data(),
             wa.data());
    Expect("HConj bx", r.bx_, wb);
    Expect("HConj ax", r.ax_, wa);
    EvkMap<word> keys;
    Evk k5;
    Host k5b, k5a;
    H.MakeKey(k5, seed + 100, k5b, k5a);
    keys.emplace(5, std::move(k5));
    std::vector<Ct> outs;
    ctx->HRotHoisted(outs, c1, keys, {5});
    or_hrot(H.op, level, b1.data(), a1.data(), k5b.data(), k5a.data(), 5, wb.data(),
            wa.data());
    Expect("HRotHoisted(5) bx", outs[0].bx_, wb);
    Expect("HRotHoisted(5) ax", outs[0].ax_, wa);
  });

  Run("MultKeyNoModDown", [&] {
    Ct acc;
    ctx->MultKeyNoModDown(acc, c1, mk);
    const size_t ts = (size_t)(np.GetNumQ() + param->alpha_) * n;
    Host wb(ts), wa(ts);
    or_mult_key_no_moddown(H.op, level, a1.data(), b1.data(), nullptr, mkb.data(),
                           mka.data(), wb.data(), wa.data());
    Expect("MultKeyNoModDown bx", acc.bx_, wb);
    Expect("MultKeyNoModDown ax", acc.ax_, wa);
    Report("MultKeyNoModDown NP", acc.GetNP().num_aux_ == param->alpha_);
  });

  // in-place MultUnsafe of a Q*P ciphertext by a Q-only plaintext and by a
  // Q-only constant: the output drops the aux limbs, so it must not reuse
  // (and reallocate) the input's buffers before reading them
  Run("In-place MultUnsafe of a Q*P ciphertext", [&] {
    const size_t ts = (size_t)(np.GetNumQ() + param->alpha_) * n;
    Host wb(ts), wa(ts);
    or_mult_key_no_moddown(H.op, level, a1.data(), b1.data(), nullptr, mkb.data(),
                           mka.data(), wb.data(), wa.data());
    Pt pt(np);
    pt.mx_.CopyFromHost(b2);
    Const k(np);
    Host kc(np.GetNumTotal());
    for (size_t t = 0; t < kc.size(); t++) kc[t] = a2[t * n + 7];  // any residues
    k.cx_.CopyFromHost(kc);
    Host pb(qs), pa(qs), cb(qs), ca(qs), kfull(qs);
    for (size_t t = 0; t < kc.size(); t++)
      std::fill(kfull.begin() + t * n, kfull.begin() + (t + 1) * n, kc[t]);
    or_mult(H.op, pidx.data(), (int)pidx.size(), wb.data(), b2.data(), pb.data());
    or_mult(H.op, pidx.data(), (int)pidx.size(), wa.data(), b2.data(), pa.data());
    or_mult(H.op, pidx.data(), (int)pidx.size(), wb.data(), kfull.data(), cb.data());
    or_mult(H.op, pidx.data(), (int)pidx.size(), wa.data(), kfull.data(), ca.data());
    Ct acc;
    ctx->MultKeyNoModDown(acc, c1, mk);
    ctx->MultUnsafe(acc, acc, pt);
    Report("in-place Ct*Pt NP", acc.GetNP() == np);
    Expect("in-place Ct*Pt bx", acc.bx_, pb);
    Expect("in-place Ct*Pt ax", acc.ax_, pa);
    ctx->MultKeyNoModDown(acc, c1, mk);
    ctx->MultUnsafe(acc, acc, k);
    Report("in-place Ct*Const NP", acc.GetNP() == np);
    Expect("in-place Ct*Const bx", acc.bx_, cb);
    Expect("in-place Ct*Const ax", acc.ax_, ca);
  });

  std::printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "OK", g_fail);
  return g_fail ? 1 : 0;
}
