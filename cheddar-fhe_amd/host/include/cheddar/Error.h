// The facade's error type: every failed check throws cheddar::Error (the
// reference asserts and exits instead, include/common/Assert.h:9-21).
#pragma once

#include <stdexcept>

namespace cheddar {

class Error : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

}  // namespace cheddar
