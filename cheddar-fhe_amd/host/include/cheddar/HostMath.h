// Host-side (CPU) arithmetic behind encoding and bootstrapping set-up: the
// special FFT of the canonical embedding, exact double -> residue
// conversion, CRT reconstruction with a small multi-word integer, and
// modular helpers. Nothing here touches the GPU.
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <utility>
#include <vector>

namespace cheddar {

using Complex = std::complex<double>;

inline bool IsPowOfTwo(long long x) { return x > 0 && (x & (x - 1)) == 0; }

inline int Log2Ceil(long long x) {
  int r = 0;
  while ((1ll << r) < x) r++;
  return r;
}

inline int Log2Floor(long long x) {
  int r = 0;
  while ((2ll << r) <= x) r++;
  return r;
}

inline int BitReverseInt(int x, int bits) {
  int r = 0;
  for (int i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
  return r;
}

// In-place bit-reversal permutation, zero-padded to the next power of two
// (include/common/CommonUtils.h:155-178).
template <typename T>
void BitReverseVector(std::vector<T> &data) {
  const int n = (int)data.size();
  const int bits = Log2Ceil(n);
  data.resize((size_t)1 << bits);
  for (int i = 0; i < (int)data.size(); i++) {
    const int j = BitReverseInt(i, bits);
    if (i < j) std::swap(data[i], data[j]);
  }
}

namespace hostmath {

inline uint32_t MulMod(uint32_t a, uint32_t b, uint32_t q) {
  return (uint32_t)((uint64_t)a * b % q);
}

inline uint32_t PowMod(uint32_t a, uint64_t e, uint32_t q) {
  uint64_t r = 1, b = a % q;
  for (; e; e >>= 1) {
    if (e & 1) r = r * b % q;
    b = b * b % q;
  }
  return (uint32_t)r;
}

inline uint32_t InvMod(uint32_t a, uint32_t q) { return PowMod(a, q - 2, q); }

// a * 2^32 mod q: the Montgomery form the GPU kernels compute with
inline uint32_t ToMontgomery(uint32_t a, uint32_t q) {
  return (uint32_t)(((uint64_t)a << 32) % q);
}

// [trunc(x)]_q in [0, q): the value of the integer that truncating x toward
// zero gives (the reference converts through BigInt(double), which
// truncates, then takes a non-negative remainder, src/core/Encode.cpp:
// 155-183). Exact for every finite double: beyond 2^63 x = m * 2^e with an
// integer mantissa m, reduced as (m mod q)(2^e mod q).
inline uint32_t DoubleToResidue(double x, uint32_t q) {
  if (!std::isfinite(x)) throw std::domain_error("non-finite value in encoding");
  const double t = std::trunc(x);
  if (std::fabs(t) < 9.2e18) {
    const int64_t v = (int64_t)t;
    int64_t r = v % (int64_t)q;
    if (r < 0) r += q;
    return (uint32_t)r;
  }
  int e;
  const double m = std::frexp(std::fabs(t), &e);  // |t| = m 2^e, m in [0.5, 1)
  const uint64_t mant = (uint64_t)std::ldexp(m, 53);
  uint64_t r = MulMod((uint32_t)(mant % q), PowMod(2, (uint64_t)(e - 53), q), q);
  if (t < 0 && r) r = q - r;
  return (uint32_t)r;
}

}  // namespace hostmath

// Unsigned multi-word integer (32-bit words, little endian), only what CRT
// reconstruction needs.
class BigUint {
 public:
  BigUint() = default;
  explicit BigUint(uint64_t v) {
    if (v) w_.push_back((uint32_t)v);
    if (v >> 32) w_.push_back((uint32_t)(v >> 32));
  }

  bool IsZero() const { return w_.empty(); }
  const std::vector<uint32_t> &Words() const { return w_; }

  void MulSmall(uint32_t m) {
    uint64_t carry = 0;
    for (auto &x : w_) {
      const uint64_t t = (uint64_t)x * m + carry;
      x = (uint32_t)t;
      carry = t >> 32;
    }
    if (carry) w_.push_back((uint32_t)carry);
    Trim();
  }

  // *this += a * m
  void AddMul(const BigUint &a, uint32_t m) {
    if (w_.size() < a.w_.size() + 1) w_.resize(a.w_.size() + 1, 0);
    uint64_t carry = 0;
    size_t i = 0;
    for (; i < a.w_.size(); i++) {
      const uint64_t t = (uint64_t)a.w_[i] * m + w_[i] + carry;
      w_[i] = (uint32_t)t;
      carry = t >> 32;
    }
    for (; carry && i < w_.size(); i++) {
      const uint64_t t = (uint64_t)w_[i] + carry;
      w_[i] = (uint32_t)t;
      carry = t >> 32;
    }
    if (carry) w_.push_back((uint32_t)carry);
    Trim();
  }

  // *this -= a, requires *this >= a
  void Sub(const BigUint &a) {
    int64_t borrow = 0;
    for (size_t i = 0; i < w_.size(); i++) {
      int64_t t = (int64_t)w_[i] - borrow - (i < a.w_.size() ? (int64_t)a.w_[i] : 0);
      borrow = t < 0;
      if (t < 0) t += (int64_t)1 << 32;
      w_[i] = (uint32_t)t;
    }
    Trim();
  }

  int Compare(const BigUint &a) const {
    if (w_.size() != a.w_.size()) return w_.size() < a.w_.size() ? -1 : 1;
    for (size_t i = w_.size(); i-- > 0;)
      if (w_[i] != a.w_[i]) return w_[i] < a.w_[i] ? -1 : 1;
    return 0;
  }

  void Half() {
    for (size_t i = 0; i < w_.size(); i++) {
      w_[i] >>= 1;
      if (i + 1 < w_.size()) w_[i] |= w_[i + 1] << 31;
    }
    Trim();
  }

  // nearest double (the top 64 bits rounded once, then scaled)
  double ToDouble() const {
    if (w_.empty()) return 0.0;
    const size_t n = w_.size();
    uint64_t top = w_[n - 1];
    int shift = 0;
    if (n >= 2) {
      top = (top << 32) | w_[n - 2];
      shift = 32 * (int)(n - 2);
    }
    const int lz = __builtin_clzll(top);
    uint64_t bits = top << lz;
    if (n >= 3 && lz) bits |= (uint64_t)w_[n - 3] >> (32 - lz);
    return std::ldexp((double)bits, shift - lz);
  }

 private:
  void Trim() {
    while (!w_.empty() && w_.back() == 0) w_.pop_back();
  }
  std::vector<uint32_t> w_;
};

// CRT reconstruction over primes q_0..q_{k-1}: given v_i = [x (Q/q_i)^-1]_{q_i},
// returns the centred x in (-Q/2, Q/2] as a double (the reference's
// PlaintextToComplexVector, src/core/Encode.cpp:184-250).
class CrtReconstructor {
 public:
  explicit CrtReconstructor(const std::vector<uint32_t> &primes) {
    q_ = BigUint(1);
    for (uint32_t p : primes) q_.MulSmall(p);
    for (size_t i = 0; i < primes.size(); i++) {
      BigUint h(1);
      for (size_t j = 0; j < primes.size(); j++)
        if (j != i) h.MulSmall(primes[j]);
      qhat_.push_back(std::move(h));
    }
    half_ = q_;
    half_.Half();
    qd_ = q_.ToDouble();
  }

  // values[i] = v_i (one per prime)
  double Centred(const uint32_t *values, size_t stride) const {
    BigUint x;
    for (size_t i = 0; i < qhat_.size(); i++) x.AddMul(qhat_[i], values[i * stride]);
    // x < (number of primes) Q: subtract k Q with k = floor(x / Q) - 1
    // estimated in doubles (never above the true quotient), then at most a
    // few single subtractions
    const double est = std::floor(x.ToDouble() / qd_) - 1.0;
    if (est >= 1.0) {
      BigUint kq = q_;
      kq.MulSmall((uint32_t)est);
      x.Sub(kq);
    }
    while (x.Compare(q_) >= 0) x.Sub(q_);
    if (x.Compare(half_) > 0) {
      BigUint neg = q_;
      neg.Sub(x);
      return -neg.ToDouble();
    }
    return x.ToDouble();
  }

 private:
  BigUint q_, half_;
  double qd_ = 0;
  std::vector<BigUint> qhat_;
};

// The special FFT pair of the CKKS canonical embedding, over num_slots
// (a power of two <= N/2) slots with the 5^j ordering. tw[i] = e^(2 pi i
// i / M), M = 2N; galois[j] = 5^j mod M. Same butterfly structure as the
// reference's Encoder (src/core/Encode.cpp:85-127), whose decoded values
// this must reproduce.
inline void SpecialIFFT(std::vector<Complex> &data, const std::vector<Complex> &tw,
                        const std::vector<int> &galois, int M) {
  const int n = (int)data.size();
  if (!IsPowOfTwo(n)) throw std::invalid_argument("Power of 2 num slots only");
  for (int stride = n / 2; stride >= 1; stride /= 2) {
    const int st8 = stride << 3, gap = M / st8;
    for (int i = 0; i < n; i += 2 * stride)
      for (int j = 0; j < stride; j++) {
        const Complex x = data[i + j] + data[i + j + stride];
        const Complex y = (data[i + j] - data[i + j + stride]) *
                          tw[(size_t)(st8 - galois[j] % st8) * gap];
        data[i + j] = x;
        data[i + j + stride] = y;
      }
  }
  BitReverseVector(data);
  for (auto &v : data) v /= (double)n;
}

inline void SpecialFFT(std::vector<Complex> &data, const std::vector<Complex> &tw,
                       const std::vector<int> &galois, int M) {
  const int n = (int)data.size();
  if (!IsPowOfTwo(n)) throw std::invalid_argument("Power of 2 num slots only");
  BitReverseVector(data);
  for (int stride = 1; stride < n; stride *= 2) {
    const int st8 = stride << 3, gap = M / st8;
    for (int i = 0; i < n; i += 2 * stride)
      for (int j = 0; j < stride; j++) {
        const Complex x = data[i + j];
        const Complex y = data[i + j + stride] * tw[(size_t)(galois[j] % st8) * gap];
        data[i + j] = x + y;
        data[i + j + stride] = x - y;
      }
  }
}

}  // namespace cheddar
