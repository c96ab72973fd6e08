// Cross-check driver (test infrastructure): exposes the reference's own
// host-side number theory (include/common/PrimeUtils.h:57-248 and
// BitReverseVector, include/common/CommonUtils.h:162-177) through a C
// interface so tests can compare the oracle's constants with the
// reference's. Built from the reference headers where they lie
// (-I /root/reference/include) into oracle/_ref/; never shipped.
#include <cstdint>
#include <vector>

#include "common/CommonUtils.h"
#include "common/PrimeUtils.h"

using namespace cheddar;

extern "C" {
uint32_t ref_find_generator(uint32_t q) { return primeutil::FindGenerator<uint32_t>(q); }
uint32_t ref_primitive_root(int m, uint32_t q) {
  return primeutil::FindPrimitiveMthRoot<uint32_t>(m, q);
}
int32_t ref_inv_mod_base(uint32_t q) { return primeutil::InvModBase<uint32_t>(q); }
uint32_t ref_to_mont(uint32_t a, uint32_t q) { return primeutil::ToMontgomery<uint32_t>(a, q); }
uint32_t ref_pow_mod(uint32_t a, int64_t e, uint32_t q) {
  return primeutil::PowMod<uint32_t>(a, e, q);
}
uint32_t ref_inv_mod(uint32_t a, uint32_t q) { return primeutil::InvMod<uint32_t>(a, q); }
uint32_t ref_mult_mod(uint32_t a, uint32_t b, uint32_t q) {
  return primeutil::MultMod<uint32_t>(a, b, q);
}
// psi_rev table exactly as PopulateTwiddleFactors builds it
// (src/core/NTT.cu:1262-1285): powers of psi, BitReverseVector, ToMontgomery
void ref_psi_rev_mont(uint32_t q, int log_n, uint32_t *out) {
  int n = 1 << log_n;
  uint32_t psi = primeutil::FindPrimitiveMthRoot<uint32_t>(2 * n, q);
  std::vector<uint32_t> v(n);
  v[0] = 1;
  for (int j = 1; j < n; j++) v[j] = primeutil::MultMod<uint32_t>(v[j - 1], psi, q);
  BitReverseVector(v);
  for (int j = 0; j < n; j++) out[j] = primeutil::ToMontgomery<uint32_t>(v[j], q);
}
}
