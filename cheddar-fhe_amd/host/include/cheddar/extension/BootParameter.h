// Bootstrapping level plan (include/extension/BootParameter.h,
// src/extension/BootParameter.cpp): CoeffToSlot from the top level, then
// EvalMod, then SlotToCoeff down to the end level. The EvalMod polynomial
// is the reference's: a degree-30 even Chebyshev series followed by three
// double-angle steps, for K = 2 * 2^3 (the coefficient table is data that
// the evaluation must share with the reference to give the same results).
#pragma once

#include <vector>

#include "cheddar/HostMath.h"

namespace cheddar {

struct BootParameter {
  BootParameter(int max_level, int num_cts_levels, int num_stc_levels,
                int log_message_ratio = 5)
      : max_level_(max_level),
        num_cts_levels_(num_cts_levels),
        num_stc_levels_(num_stc_levels),
        log_message_ratio_(log_message_ratio),
        mod_coefficients_{
            0.12517186708929745802,    0.0, 0.2894364973331168731,      0.0,
            0.36272381596524499154,    0.0, 0.3011054704600794278,      0.0,
            -0.10550875667295944105,   0.0, -0.43588877795190139706,    0.0,
            0.37482647434055190702,    0.0, -0.14821069913569220404,    0.0,
            0.03665437786710548091,    0.0, -0.0063882548960017121343,  0.0,
            0.00083684232451067872756, 0.0, -8.6443599931576702305e-05, 0.0,
            7.0966437900548814324e-06, 0.0, -5.228015817181348194e-07,  0.0,
            2.2714690137973883081e-08, 0.0, -2.3761936068138980797e-09},
        num_double_angle_(3),
        initial_K_(2) {}

  const int max_level_;
  const int num_cts_levels_;
  const int num_stc_levels_;
  const int log_message_ratio_;
  const std::vector<double> mod_coefficients_;
  const int num_double_angle_;
  const int initial_K_;

  int GetNumEvalModLevels() const {
    return Log2Ceil((long long)mod_coefficients_.size()) + num_double_angle_;
  }
  int GetMaxLevel() const { return max_level_; }
  int GetCtSStartLevel() const { return max_level_; }
  int GetEvalModStartLevel() const { return max_level_ - num_cts_levels_; }
  int GetStCStartLevel() const { return GetEvalModStartLevel() - GetNumEvalModLevels(); }
  int GetStartLevel() const { return max_level_; }
  int GetEndLevel() const { return GetStCStartLevel() - num_stc_levels_; }
};

}  // namespace cheddar
