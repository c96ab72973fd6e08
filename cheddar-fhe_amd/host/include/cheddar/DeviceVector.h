// C++ host mirror of the reference's DeviceVector (include/core/
// DeviceVector.h): an owning, move-only array of words in GPU memory.
// Allocation uses hipMalloc; errors throw cheddar::Error (the reference
// asserts and exits instead, include/common/Assert.h:9-21).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace cheddar {

class Error : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

inline void HipCheck(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename word>
class DeviceVector {
 public:
  DeviceVector() = default;
  explicit DeviceVector(size_t size) { resize(size); }
  ~DeviceVector() { release(); }

  DeviceVector(const DeviceVector &) = delete;
  DeviceVector &operator=(const DeviceVector &) = delete;
  DeviceVector(DeviceVector &&o) noexcept
      : ptr_(std::exchange(o.ptr_, nullptr)), size_(std::exchange(o.size_, 0)) {}
  DeviceVector &operator=(DeviceVector &&o) noexcept {
    if (this != &o) {
      release();
      ptr_ = std::exchange(o.ptr_, nullptr);
      size_ = std::exchange(o.size_, 0);
    }
    return *this;
  }

  size_t size() const { return size_; }
  word *data() { return ptr_; }
  const word *data() const { return ptr_; }

  // Reallocates when the size changes; the leading min(old, new) words are
  // kept.
  void resize(size_t size) {
    if (size == size_) return;
    word *p = nullptr;
    if (size) HipCheck(hipMalloc((void **)&p, size * sizeof(word)), "hipMalloc");
    if (ptr_ && p && size_)
      HipCheck(hipMemcpy(p, ptr_, std::min(size, size_) * sizeof(word),
                         hipMemcpyDeviceToDevice),
               "hipMemcpy");
    release();
    ptr_ = p;
    size_ = size;
  }

  void CopyFromHost(const std::vector<word> &h) {
    resize(h.size());
    if (!h.empty())
      HipCheck(hipMemcpy(ptr_, h.data(), h.size() * sizeof(word),
                         hipMemcpyHostToDevice),
               "hipMemcpy H2D");
  }

  std::vector<word> CopyToHost() const {
    std::vector<word> h(size_);
    if (size_)
      HipCheck(hipMemcpy(h.data(), ptr_, size_ * sizeof(word),
                         hipMemcpyDeviceToHost),
               "hipMemcpy D2H");
    return h;
  }

 private:
  void release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    size_ = 0;
  }
  word *ptr_ = nullptr;
  size_t size_ = 0;
};

}  // namespace cheddar
