// Encrypted-semantics test of the C++ host facade on the GPU, the checks of
// the reference's unittest/BasicTest.cpp: encode / decode, encrypt /
// decrypt, ciphertext +- ciphertext / plaintext / constant, negation,
// products by plaintexts, constants and the imaginary unit, HMult (merged
// and separate rescale), HRot and HConj, each decrypted and decoded and
// compared with the same operation on the complex messages (max error
// 1e-3, as the reference's CompareMessages). Keys come from the host
// UserInterface (seeded).
//
//   test_basic <params.json> [dense_hamming_weight]
//
// Prints one line per case (max error and precision bits) and exits
// non-zero if any case fails.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

#include "cheddar/Context.h"
#include "cheddar/ParameterJson.h"
#include "cheddar/UserInterface.h"

using namespace cheddar;
using word = uint32_t;
using Ct = Ciphertext<word>;
using Pt = Plaintext<word>;
using Const = Constant<word>;
using Msg = std::vector<Complex>;

namespace {

int g_fail = 0;
constexpr double kMaxError = 1e-3;

struct Bed {
  std::unique_ptr<Parameter<word>> param;
  std::shared_ptr<Context<word>> ctx;
  std::unique_ptr<UserInterface<word>> ui;
  Random rng{0x4348454444415204ull};

  double Scale(int level) const {
    return level <= param->default_encryption_level_ ? param->GetScale(level)
                                                     : param->GetRescalePrimeProd(level);
  }
  Msg RandomMsg(int n) {
    Msg m(n);
    rng.SampleUniformComplex(m.data(), n, -1.0, 1.0);
    return m;
  }
  void Encode(Pt &pt, const Msg &m, int level) {
    ctx->encoder_.Encode(pt, level, Scale(level), m);
  }
  void EncodeAndEncrypt(Ct &ct, const Msg &m, int level) {
    Pt pt;
    Encode(pt, m, level);
    ui->Encrypt(ct, pt);
  }
  Msg DecryptAndDecode(const Ct &ct) {
    Pt pt;
    ui->Decrypt(pt, ct);
    Msg m;
    ctx->encoder_.Decode(m, pt);
    return m;
  }
};

void Compare(const std::string &name, const Msg &want, const Msg &got,
             double max_error = kMaxError) {
  double worst = 0;
  bool ok = want.size() == got.size();
  for (size_t i = 0; ok && i < want.size(); i++) {
    const double e = std::abs(want[i] - got[i]);
    if (!(e <= worst)) worst = e;  // NaN propagates
  }
  ok = ok && worst < max_error;
  std::printf("%s %s: max error %.3e (%.1f bits)\n", ok ? "PASS" : "FAIL", name.c_str(),
              worst, -std::log2(worst));
  std::fflush(stdout);
  if (!ok) g_fail++;
}

template <typename F>
void Run(const std::string &name, F f) {
  try {
    f();
  } catch (const std::exception &e) {
    std::printf("FAIL %s: %s\n", name.c_str(), e.what());
    std::fflush(stdout);
    g_fail++;
  }
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <params.json> [dense_hamming_weight]\n", argv[0]);
    return 2;
  }
  Bed b;
  b.param = ParameterFromJsonFile(argv[1]);
  if (argc > 2) {
    b.param->SetSparseHammingWeight(std::min(b.param->GetSparseHammingWeight(),
                                             std::atoi(argv[2])));
    b.param->SetDenseHammingWeight(std::atoi(argv[2]));
  }
  b.ctx = Context<word>::Create(*b.param);
  b.ui = std::make_unique<UserInterface<word>>(b.ctx, 0x4348454444415205ull);
  const Parameter<word> &P = *b.param;
  const int slots = P.degree_ / 2;
  const int top = P.max_level_;
  // a level can hold a message when its modulus exceeds the scale by a
  // margin (the 24-prime chain's level 0 is one 31-bit prime under a 2^40
  // scale, which no message fits)
  auto usable = [&](int level) {
    double bits = 0;
    for (word q : P.GetPrimeVector(P.LevelToNP(level))) bits += std::log2((double)q);
    return bits >= std::log2(b.Scale(level)) + 10;
  };
  int low = 0;
  while (low < top && !usable(low)) low++;
  std::vector<int> levels = {low, low + 1, (low + top) / 2, top};
  std::printf("params %s: N=%d, %d levels, dnum %d, alpha %d, h %d\n", argv[1],
              P.degree_, top + 1, P.dnum_, P.alpha_, P.GetDenseHammingWeight());

  for (int level : levels) {
    const std::string at = " at level " + std::to_string(level);
    Run("EncodeDecode" + at, [&] {
      const Msg m = b.RandomMsg(slots);
      Pt pt;
      b.Encode(pt, m, level);
      Msg r;
      b.ctx->encoder_.Decode(r, pt);
      Compare("EncodeDecode" + at, m, r);
    });
    Run("EncodeEncryptDecryptDecode" + at, [&] {
      const Msg m = b.RandomMsg(slots);
      Ct ct;
      b.EncodeAndEncrypt(ct, m, level);
      Compare("EncodeEncryptDecryptDecode" + at, m, b.DecryptAndDecode(ct));
    });
    Run("CtAddCt / CtSubCt / NegCt" + at, [&] {
      const Msg m1 = b.RandomMsg(slots), m2 = b.RandomMsg(slots);
      Ct c1, c2, r;
      b.EncodeAndEncrypt(c1, m1, level);
      b.EncodeAndEncrypt(c2, m2, level);
      Msg add(slots), sub(slots), neg(slots);
      for (int i = 0; i < slots; i++) {
        add[i] = m1[i] + m2[i];
        sub[i] = m1[i] - m2[i];
        neg[i] = -m1[i];
      }
      b.ctx->Add(r, c1, c2);
      Compare("CtAddCt" + at, add, b.DecryptAndDecode(r));
      b.ctx->Sub(r, c1, c2);
      Compare("CtSubCt" + at, sub, b.DecryptAndDecode(r));
      b.ctx->Neg(r, c1);
      Compare("NegCt" + at, neg, b.DecryptAndDecode(r));
    });
    Run("CtAddPt / CtSubPt" + at, [&] {
      const Msg m1 = b.RandomMsg(slots), m2 = b.RandomMsg(slots);
      Ct c1, r;
      Pt p2;
      b.EncodeAndEncrypt(c1, m1, level);
      b.Encode(p2, m2, level);
      Msg add(slots), sub(slots), rsub(slots);
      for (int i = 0; i < slots; i++) {
        add[i] = m1[i] + m2[i];
        sub[i] = m1[i] - m2[i];
        rsub[i] = m2[i] - m1[i];
      }
      b.ctx->Add(r, c1, p2);
      Compare("CtAddPt" + at, add, b.DecryptAndDecode(r));
      b.ctx->Sub(r, c1, p2);
      Compare("CtSubPt" + at, sub, b.DecryptAndDecode(r));
      b.ctx->Sub(r, p2, c1);
      Compare("PtSubCt" + at, rsub, b.DecryptAndDecode(r));
    });
    Run("CtAddConst / CtSubConst" + at, [&] {
      const Msg m1 = b.RandomMsg(slots);
      const double c = b.RandomMsg(1)[0].real();
      Ct c1, r;
      Const k;
      b.EncodeAndEncrypt(c1, m1, level);
      b.ctx->encoder_.EncodeConstant(k, level, b.Scale(level), c);
      Msg add(slots), sub(slots), rsub(slots);
      for (int i = 0; i < slots; i++) {
        add[i] = m1[i] + c;
        sub[i] = m1[i] - c;
        rsub[i] = c - m1[i];
      }
      b.ctx->Add(r, c1, k);
      Compare("CtAddConst" + at, add, b.DecryptAndDecode(r));
      b.ctx->Sub(r, c1, k);
      Compare("CtSubConst" + at, sub, b.DecryptAndDecode(r));
      b.ctx->Sub(r, k, c1);
      Compare("ConstSubCt" + at, rsub, b.DecryptAndDecode(r));
    });
    Run("CtMultImaginaryUnit" + at, [&] {
      const Msg m1 = b.RandomMsg(slots);
      Ct c1, r;
      b.EncodeAndEncrypt(c1, m1, level);
      Msg want(slots);
      for (int i = 0; i < slots; i++) want[i] = Complex(-m1[i].imag(), m1[i].real());
      b.ctx->MultImaginaryUnit(r, c1);
      Compare("CtMultImaginaryUnit" + at, want, b.DecryptAndDecode(r));
    });
    if (level >= 1 && usable(level - 1)) {
      Run("CtMultPt / CtMultConst" + at, [&] {
        const Msg m1 = b.RandomMsg(slots), m2 = b.RandomMsg(slots);
        const double c = m2[0].real();
        Ct c1, r;
        Pt p2;
        Const k;
        b.EncodeAndEncrypt(c1, m1, level);
        b.Encode(p2, m2, level);
        b.ctx->encoder_.EncodeConstant(k, level, b.Scale(level), c);
        Msg mp(slots), mc(slots);
        for (int i = 0; i < slots; i++) {
          mp[i] = m1[i] * m2[i];
          mc[i] = m1[i] * c;
        }
        b.ctx->Mult(r, c1, p2);
        Compare("CtMultPt" + at, mp, b.DecryptAndDecode(r));
        b.ctx->Mult(r, c1, k);
        Compare("CtMultConst" + at, mc, b.DecryptAndDecode(r));
      });
      Run("HMult" + at, [&] {
        const Msg m1 = b.RandomMsg(slots), m2 = b.RandomMsg(slots);
        Ct c1, c2, r, t;
        b.EncodeAndEncrypt(c1, m1, level);
        b.EncodeAndEncrypt(c2, m2, level);
        Msg want(slots);
        for (int i = 0; i < slots; i++) want[i] = m1[i] * m2[i];
        b.ctx->HMult(r, c1, c2, b.ui->GetMultiplicationKey(), true);
        Compare("HMult (merged rescale)" + at, want, b.DecryptAndDecode(r));
        b.ctx->HMult(t, c1, c2, b.ui->GetMultiplicationKey(), false);
        Compare("HMult (relinearize)" + at, want, b.DecryptAndDecode(t));
        b.ctx->Rescale(r, t);
        Compare("Rescale" + at, want, b.DecryptAndDecode(r));
      });
    }
  }

  const int rot = 1234;
  b.ui->PrepareRotationKey(rot, top);
  for (int level : levels) {
    const std::string at = " at level " + std::to_string(level);
    Run("HRot" + at, [&] {
      const Msg m1 = b.RandomMsg(slots);
      Ct c1, r;
      b.EncodeAndEncrypt(c1, m1, level);
      Msg want(slots);
      for (int i = 0; i < slots; i++) want[i] = m1[(i + rot) % slots];
      b.ctx->HRot(r, c1, b.ui->GetRotationKey(rot), rot);
      Compare("HRot" + at, want, b.DecryptAndDecode(r));
    });
    Run("HConj" + at, [&] {
      const Msg m1 = b.RandomMsg(slots);
      Ct c1, r;
      b.EncodeAndEncrypt(c1, m1, level);
      Msg want(slots);
      for (int i = 0; i < slots; i++) want[i] = std::conj(m1[i]);
      b.ctx->HConj(r, c1, b.ui->GetConjugationKey());
      Compare("HConj" + at, want, b.DecryptAndDecode(r));
    });
  }
  if (top >= 2) {
    Run("LevelDown", [&] {
      const int from = std::min(top, P.default_encryption_level_);
      const Msg m1 = b.RandomMsg(slots);
      Ct c1, r;
      b.EncodeAndEncrypt(c1, m1, from);
      b.ctx->LevelDown(r, c1, low);
      Compare("LevelDown " + std::to_string(from) + " -> " + std::to_string(low), m1,
              b.DecryptAndDecode(r));
    });
  }
  std::printf("OK: %d failure(s)\n", g_fail);
  return g_fail ? 1 : 0;
}
