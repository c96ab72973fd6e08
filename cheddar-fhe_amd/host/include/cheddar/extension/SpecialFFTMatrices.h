// The butterfly stages of the encoder's special FFT as sparse (three-
// diagonal) matrices (EvalSpecialFFT::PopulatePlainMatrices,
// src/extension/EvalSpecialFFT.cpp:66-160). With n slots, M = 2N, tw[i] =
// e^(2 pi i i / M) and galois[j] = 5^j mod M:
//   fft[s]  (stride 2^s):  (x, y) -> (x + w y, x - w y)
//   ifft[s] (stride n / 2^(s+1)): (x, y) -> (x + y, (x - y) w^-1)
// so that applying fft[0], fft[1], ... to v gives SpecialFFT(bitrev(v)) and
// applying ifft[0], ifft[1], ... gives n bitrev(SpecialIFFT(v)).
#pragma once

#include <vector>

#include "cheddar/HostMath.h"
#include "cheddar/extension/StripedMatrix.h"

namespace cheddar {

inline void SpecialFFTStageMatrices(int num_slots, int M, const std::vector<int> &galois,
                                    const std::vector<Complex> &tw,
                                    std::vector<StripedMatrix> &fft,
                                    std::vector<StripedMatrix> &ifft) {
  const int stages = Log2Ceil(num_slots);
  fft.assign(stages, StripedMatrix());
  ifft.assign(stages, StripedMatrix());
  for (int i = 0; i < stages; i++) {
    const int stride = 1 << i, group = 2 * stride, st8 = stride << 3, gap = M / st8;
    StripedMatrix &f = fft[i];
    StripedMatrix &g = ifft[stages - 1 - i];
    f = StripedMatrix(num_slots, num_slots);
    g = StripedMatrix(num_slots, num_slots);
    const bool last = i == stages - 1;
    for (StripedMatrix *m : {&f, &g}) {
      m->emplace(0, std::vector<Complex>(num_slots));
      m->emplace(stride, std::vector<Complex>(num_slots));
      if (!last) m->emplace(num_slots - stride, std::vector<Complex>(num_slots));
    }
    // for the last stage stride == num_slots - stride: one shared diagonal
    auto &f0 = f[0], &fp = f[stride], &fm = f[num_slots - stride];
    auto &g0 = g[0], &gp = g[stride], &gm = g[num_slots - stride];
    for (int j = 0; j < stride; j++) {
      const int gal = galois[j] % st8;
      const Complex wf = tw[(size_t)gal * gap];
      const Complex wi = tw[(size_t)(st8 - gal) * gap];
      f0[j] = 1;
      fp[j] = wf;
      fm[j + stride] = 1;
      f0[j + stride] = -wf;
      g0[j] = 1;
      gp[j] = 1;
      gm[j + stride] = wi;
      g0[j + stride] = -wi;
    }
    // every group of 2 stride entries repeats the first
    for (int start = group; start < num_slots; start *= 2)
      for (auto *d : {&f0, &fp, &fm, &g0, &gp, &gm}) {
        if (last && (d == &fm || d == &gm)) continue;
        std::copy(d->begin(), d->begin() + start, d->begin() + start);
      }
  }
}

// y = M x for a striped matrix: y[k] = sum_i diag_i[k] x[(k + i) mod n]
inline std::vector<Complex> StripedApply(const StripedMatrix &m, const std::vector<Complex> &x) {
  const int n = (int)x.size();
  std::vector<Complex> y(n, Complex(0));
  for (const auto &[i, d] : m)
    for (int k = 0; k < n; k++) y[k] += d[k] * x[(k + i) % n];
  return y;
}

}  // namespace cheddar
