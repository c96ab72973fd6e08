// C++ host mirror of the reference's NPInfo (include/core/NPInfo.h): the
// number of main, terminal and auxiliary primes a polynomial is held in.
#pragma once

namespace cheddar {

struct NPInfo {
  int num_main_ = 0;
  int num_ter_ = 0;
  int num_aux_ = 0;

  explicit NPInfo(int num_main = 0, int num_ter = 0, int num_aux = 0)
      : num_main_(num_main), num_ter_(num_ter), num_aux_(num_aux) {}

  int GetNumQ() const { return num_main_ + num_ter_; }
  int GetNumTotal() const { return num_main_ + num_ter_ + num_aux_; }

  bool operator==(const NPInfo &o) const {
    return num_main_ == o.num_main_ && num_ter_ == o.num_ter_ &&
           num_aux_ == o.num_aux_;
  }
  bool operator!=(const NPInfo &o) const { return !(*this == o); }
  bool IsSubsetOf(const NPInfo &o) const {
    return num_main_ <= o.num_main_ && num_ter_ <= o.num_ter_ &&
           num_aux_ <= o.num_aux_;
  }
  bool IsSupersetOf(const NPInfo &o) const { return o.IsSubsetOf(*this); }
};

}  // namespace cheddar
