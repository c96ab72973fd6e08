// C++ host mirror of the reference's DeviceVector (include/core/
// DeviceVector.h): an owning, move-only array of words in GPU memory.
// Errors throw cheddar::Error (the reference asserts and exits instead,
// include/common/Assert.h:9-21).
//
// Memory comes from a caching pool (the reference allocates through an RMM
// binning pool over stream-ordered allocation, src/core/MemoryPool.cpp):
// blocks are kept per exact size and handed out again, so evaluation does
// not call hipMalloc / hipFree (which synchronises the device) once the
// working set is resident. Polynomial sizes are multiples of one limb, so
// few distinct sizes occur. A block is reused in issue order: all work on
// pooled memory goes to one stream (the Context's), or is synchronous.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "cheddar/Error.h"

namespace cheddar {

inline void HipCheck(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

class DevicePool {
 public:
  // Blocks freed while a scope is active stay with the scope (and are handed
  // out again only inside it) until the scope is released: a captured HIP
  // graph keeps using the temporaries it was recorded with, so they must
  // not go to other owners while the graph lives.
  struct Scope {
    std::unordered_map<size_t, std::vector<void *>> free;
  };

  static DevicePool &Get() {
    static DevicePool *pool = new DevicePool;  // outlives every DeviceVector
    return *pool;
  }
  void BeginScope(Scope *s) {
    std::lock_guard<std::mutex> g(m_);
    scope_ = s;
  }
  void EndScope() {
    std::lock_guard<std::mutex> g(m_);
    scope_ = nullptr;
  }
  Scope *CurrentScope() {
    std::lock_guard<std::mutex> g(m_);
    return scope_;
  }
  // move s's blocks into scope `into`, or back to the pool when into is null
  void MergeScope(Scope *s, Scope *into) {
    if (!into) return ReleaseScope(s);
    std::lock_guard<std::mutex> g(m_);
    for (auto &kv : s->free) {
      auto &dst = into->free[kv.first];
      dst.insert(dst.end(), kv.second.begin(), kv.second.end());
    }
    s->free.clear();
  }
  // give a finished scope's blocks back to the pool
  void ReleaseScope(Scope *s) {
    std::lock_guard<std::mutex> g(m_);
    for (auto &kv : s->free)
      for (void *p : kv.second) {
        free_[kv.first].push_back(p);
        cached_ += kv.first;
      }
    s->free.clear();
  }
  void *Alloc(size_t bytes) {
    {
      std::lock_guard<std::mutex> g(m_);
      if (scope_) {
        auto it = scope_->free.find(bytes);
        if (it != scope_->free.end() && !it->second.empty()) {
          void *p = it->second.back();
          it->second.pop_back();
          return p;
        }
      }
      auto it = free_.find(bytes);
      if (it != free_.end() && !it->second.empty()) {
        void *p = it->second.back();
        it->second.pop_back();
        cached_ -= bytes;
        return p;
      }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      Trim();  // give cached blocks back and try once more
      HipCheck(hipMalloc(&p, bytes), "hipMalloc");
    }
    return p;
  }
  void Free(void *p, size_t bytes) {
    std::lock_guard<std::mutex> g(m_);
    if (scope_) {
      scope_->free[bytes].push_back(p);
      return;
    }
    free_[bytes].push_back(p);
    cached_ += bytes;
  }
  // return every cached block to the driver
  void Trim() {
    std::lock_guard<std::mutex> g(m_);
    for (auto &kv : free_)
      for (void *p : kv.second) (void)hipFree(p);
    free_.clear();
    cached_ = 0;
  }
  size_t CachedBytes() const { return cached_; }

 private:
  std::mutex m_;
  std::unordered_map<size_t, std::vector<void *>> free_;
  size_t cached_ = 0;
  Scope *scope_ = nullptr;
};

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
  // kept. The copy waits for all device work first: the old contents may
  // still be written by kernels on any stream (a plain hipMemcpy is not
  // ordered against non-blocking streams).
  void resize(size_t size) {
    if (size == size_) return;
    word *p = size ? (word *)DevicePool::Get().Alloc(size * sizeof(word)) : nullptr;
    if (ptr_ && p && size_) {
      HipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
      HipCheck(hipMemcpy(p, ptr_, std::min(size, size_) * sizeof(word),
                         hipMemcpyDeviceToDevice),
               "hipMemcpy");
    }
    release();
    ptr_ = p;
    size_ = size;
  }

  // Reallocates when the size changes, without keeping the contents (for
  // outputs that are about to be overwritten).
  void reset(size_t size) {
    if (size == size_) return;
    release();
    ptr_ = size ? (word *)DevicePool::Get().Alloc(size * sizeof(word)) : nullptr;
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
    if (ptr_) DevicePool::Get().Free(ptr_, size_ * sizeof(word));
    ptr_ = nullptr;
    size_ = 0;
  }
  word *ptr_ = nullptr;
  size_t size_ = 0;
};

}  // namespace cheddar
