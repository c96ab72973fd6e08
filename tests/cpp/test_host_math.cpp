// CPU test of the host-side arithmetic behind encoding (cheddar/HostMath.h),
// against what the mathematics fixes:
// - the special IFFT then FFT is the identity;
// - the encoded polynomial evaluated at zeta^(5^j) (zeta = e^(2 pi i / 2N))
//   returns slot j (the canonical embedding, brute force at N = 32);
// - DoubleToResidue equals [trunc(x)]_q computed in 128-bit integers, for
//   values up to 2^120;
// - CRT reconstruction returns the centred integer (128-bit reference);
// - striped (diagonal-stored) matrix products equal dense products;
// - the special-FFT stage matrices of bootstrapping compose to the encoder's
//   special FFT and inverse (with the bit reversal and the 1/n they leave
//   out).
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "cheddar/HostMath.h"
#include "cheddar/extension/SpecialFFTMatrices.h"
#include "cheddar/extension/StripedMatrix.h"

using namespace cheddar;

static int fails = 0;
static void Check(bool ok, const char *what) {
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", what);
  if (!ok) fails++;
}

int main() {
  std::mt19937_64 g(0x4348454444415206ull);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  for (int logn : {5, 12, 16}) {
    const int n = 1 << logn, M = 2 * n, slots = n / 2;
    std::vector<Complex> tw(M);
    for (int i = 0; i < M; i++) tw[i] = std::polar(1.0, 2.0 * M_PI * i / M);
    std::vector<int> gal(slots + 1, 1);
    for (int i = 1; i <= slots; i++) gal[i] = (int)((long long)gal[i - 1] * 5 % M);
    std::vector<Complex> m(slots);
    for (auto &v : m) v = Complex(u(g), u(g));
    std::vector<Complex> c = m;
    SpecialIFFT(c, tw, gal, M);
    std::vector<Complex> back = c;
    SpecialFFT(back, tw, gal, M);
    double e = 0;
    for (int i = 0; i < slots; i++) e = std::fmax(e, std::abs(back[i] - m[i]));
    Check(e < 1e-12, ("IFFT then FFT is the identity, N = 2^" + std::to_string(logn)).c_str());
    if (logn == 5) {
      // p(X) = sum_i Re(c_i) X^i + Im(c_i) X^(i + N/2); p(zeta^(5^j)) = m_j
      double worst = 0;
      for (int j = 0; j < slots; j++) {
        const Complex z = std::polar(1.0, 2.0 * M_PI * gal[j] / M);
        Complex acc = 0;
        for (int i = 0; i < slots; i++)
          acc += c[i].real() * std::pow(z, i) + c[i].imag() * std::pow(z, i + slots);
        worst = std::fmax(worst, std::abs(acc - m[j]));
      }
      Check(worst < 1e-12, "encoded polynomial evaluates to the slots (canonical embedding)");
    }
  }

  const uint32_t qs[3] = {2147352577u, 30539777u, 1083703297u};
  bool ok = true;
  for (int t = 0; t < 20000 && ok; t++) {
    const int e = (int)(g() % 120);
    const double x = std::ldexp(u(g), e);
    const double tr = std::trunc(x);
    for (uint32_t q : qs) {
      // exact [trunc(x)]_q with 128-bit integers: tr = m 2^k, |tr| < 2^120
      int k;
      const double fr = std::frexp(std::fabs(tr), &k);
      __int128 v = (__int128)(uint64_t)std::ldexp(fr, 53);
      int sh = k - 53;
      if (sh >= 0) v <<= sh; else v >>= -sh;
      long long r = (long long)(v % q);
      if (tr < 0 && r) r = q - r;
      if ((uint32_t)r != hostmath::DoubleToResidue(x, q)) ok = false;
    }
  }
  Check(ok, "DoubleToResidue equals the truncated integer mod q (128-bit reference)");

  std::vector<uint32_t> primes(qs, qs + 3);
  CrtReconstructor crt(primes);
  __int128 Q = 1;
  for (uint32_t q : qs) Q *= q;
  ok = true;
  for (int t = 0; t < 20000 && ok; t++) {
    __int128 x = (__int128)(((unsigned __int128)g() << 64 | g()) % (unsigned __int128)Q);
    x -= Q / 2;  // (-Q/2, Q/2]
    if (t == 0) x = Q / 2;
    if (t == 1) x = -(Q / 2) + 1;
    if (t == 2) x = 0;
    uint32_t v[3];
    for (int i = 0; i < 3; i++) {
      const __int128 qh = Q / qs[i];
      const uint32_t inv = hostmath::InvMod((uint32_t)(qh % qs[i]), qs[i]);
      long long xr = (long long)(x % qs[i]);
      if (xr < 0) xr += qs[i];
      v[i] = hostmath::MulMod((uint32_t)xr, inv, qs[i]);
    }
    const double want = (double)x, got = crt.Centred(v, 1);
    if (std::fabs(got - want) > 1e-15 * std::fabs(want)) ok = false;
  }
  Check(ok, "CRT reconstruction returns the centred integer");

  {
    // striped product vs dense product
    const int n = 16;
    auto random_striped = [&](std::initializer_list<int> diags) {
      StripedMatrix m(n, n);
      for (int d : diags) {
        std::vector<Complex> v(n);
        for (auto &e : v) e = Complex(u(g), u(g));
        m.emplace(d, v);
      }
      return m;
    };
    auto dense = [&](const StripedMatrix &m) {
      std::vector<std::vector<Complex>> D(n, std::vector<Complex>(n));
      for (const auto &[i, d] : m)
        for (int k = 0; k < n; k++) D[k][(k + i) % n] += d[k];
      return D;
    };
    const StripedMatrix a = random_striped({0, 3, 15}), b = random_striped({1, 2, 8});
    const auto C = dense(StripedMatrix::Mult(a, b)), A = dense(a), B = dense(b);
    double e = 0;
    for (int r = 0; r < n; r++)
      for (int c = 0; c < n; c++) {
        Complex acc = 0;
        for (int t = 0; t < n; t++) acc += A[r][t] * B[t][c];
        e = std::fmax(e, std::abs(acc - C[r][c]));
      }
    Check(e < 1e-12, "striped matrix product equals the dense product");
  }

  {
    // stage matrices compose to the special FFT pair
    const int logn = 12, n = 1 << logn, M = 2 * n, slots = n / 2;
    std::vector<Complex> tw(M);
    for (int i = 0; i < M; i++) tw[i] = std::polar(1.0, 2.0 * M_PI * i / M);
    std::vector<int> gal(slots + 1, 1);
    for (int i = 1; i <= slots; i++) gal[i] = (int)((long long)gal[i - 1] * 5 % M);
    std::vector<StripedMatrix> fft, ifft;
    SpecialFFTStageMatrices(slots, M, gal, tw, fft, ifft);
    std::vector<Complex> v(slots);
    for (auto &x : v) x = Complex(u(g), u(g));
    std::vector<Complex> y = v;
    for (const auto &m : fft) y = StripedApply(m, y);
    std::vector<Complex> want = v;
    BitReverseVector(want);
    SpecialFFT(want, tw, gal, M);
    double e = 0;
    for (int i = 0; i < slots; i++) e = std::fmax(e, std::abs(y[i] - want[i]));
    Check(e < 1e-9, "forward stage matrices compose to SpecialFFT(bitrev(v))");
    y = v;
    for (const auto &m : ifft) y = StripedApply(m, y);
    want = v;
    SpecialIFFT(want, tw, gal, M);
    BitReverseVector(want);
    e = 0;
    for (int i = 0; i < slots; i++) e = std::fmax(e, std::abs(y[i] - (double)slots * want[i]));
    Check(e < 1e-9, "inverse stage matrices compose to n bitrev(SpecialIFFT(v))");
  }
  std::printf("OK: %d failure(s)\n", fails);
  return fails ? 1 : 0;
}
