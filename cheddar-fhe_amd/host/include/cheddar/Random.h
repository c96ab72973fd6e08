// Seeded sampling for test-side key generation and messages, with the
// reference's sampling routines (include/Random.h): uniform words in a
// range, sampling without replacement, uniform reals / complex numbers and
// a rounded normal clamped at 6 sigma. The reference draws from a
// random_device-seeded mt19937_64; here every Random object is seeded
// explicitly (SplitMix64 -> xoshiro256**), so runs are reproducible. Not a
// cryptographic generator: for tests only, as the reference says of its own.
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

namespace cheddar {

class Random {
 public:
  explicit Random(uint64_t seed) {
    uint64_t z = seed;
    for (auto &w : s_) w = SplitMix(z);
  }

  uint64_t Next() {
    const uint64_t r = Rotl(s_[1] * 5, 7) * 9;
    const uint64_t t = s_[1] << 17;
    s_[2] ^= s_[0];
    s_[3] ^= s_[1];
    s_[1] ^= s_[2];
    s_[0] ^= s_[3];
    s_[2] ^= t;
    s_[3] = Rotl(s_[3], 45);
    return r;
  }

  // uniform integer in [lo, hi] by rejection (no modulo bias)
  uint64_t UniformInt(uint64_t lo, uint64_t hi) {
    const uint64_t span = hi - lo + 1;
    if (span == 0) return Next();
    const uint64_t limit = UINT64_MAX - UINT64_MAX % span;
    uint64_t v;
    do v = Next(); while (v >= limit);
    return lo + v % span;
  }

  // uniform double in [0, 1)
  double UniformReal() { return (double)(Next() >> 11) * 0x1.0p-53; }

  template <typename word>
  void SampleUniformWord(word *ptr, int length, word range_min, word range_max) {
    for (int i = 0; i < length; i++) ptr[i] = (word)UniformInt(range_min, range_max);
  }

  // `length` distinct integers of [range_min, range_max] (partial
  // Fisher-Yates shuffle)
  template <typename T>
  void SampleWithoutReplacement(T *ptr, int length, int range_min, int range_max) {
    std::vector<int> v(range_max - range_min + 1);
    for (size_t i = 0; i < v.size(); i++) v[i] = range_min + (int)i;
    for (int i = 0; i < length; i++) {
      const size_t j = i + UniformInt(0, v.size() - 1 - i);
      std::swap(v[i], v[j]);
      ptr[i] = (T)v[i];
    }
  }

  template <typename T>
  void SampleUniformReal(T *ptr, int length, double range_min, double range_max) {
    for (int i = 0; i < length; i++)
      ptr[i] = (T)(range_min + (range_max - range_min) * UniformReal());
  }

  void SampleUniformComplex(std::complex<double> *ptr, int length,
                            double range_min, double range_max) {
    for (int i = 0; i < length; i++) {
      const double re = range_min + (range_max - range_min) * UniformReal();
      const double im = range_min + (range_max - range_min) * UniformReal();
      ptr[i] = {re, im};
    }
  }

  // rounded N(mean, stddev^2), clamped at +-max_sigma stddev (Box-Muller)
  template <typename T>
  void SampleDiscreteNormal(T *ptr, int length, double mean, double stddev,
                            double max_sigma = 6) {
    for (int i = 0; i < length; i++) {
      double u1;
      do u1 = UniformReal(); while (u1 <= 0.0);
      const double u2 = UniformReal();
      double v = stddev * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
      v = std::fmax(-max_sigma * stddev, std::fmin(max_sigma * stddev, v));
      ptr[i] = (T)std::llround(v + mean);
    }
  }

 private:
  static uint64_t Rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  static uint64_t SplitMix(uint64_t &z) {
    uint64_t r = (z += 0x9E3779B97F4A7C15ull);
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
    return r ^ (r >> 31);
  }
  uint64_t s_[4];
};

}  // namespace cheddar
