// Owning buffers for the facade's users and tests: device_vector (device
// memory) and pinned_vector (hipHostMalloc'd host memory), the roles of
// thrust::device_vector and the reference's pinned_vector
// (include/rtspatial/utils/util.h:106-113).
#ifndef RTSPATIAL_AMD_UTILS_MEMORY_H
#define RTSPATIAL_AMD_UTILS_MEMORY_H

#include <hip/hip_runtime_api.h>

#include <cstring>
#include <vector>

#include "rtspatial/utils/exception.h"

namespace rtspatial {

template <typename T>
class device_vector {
 public:
  device_vector() = default;
  explicit device_vector(size_t n) { resize(n); }
  device_vector(const std::vector<T>& host) { assign(host.data(), host.size()); }
  ~device_vector() { release(); }
  device_vector(const device_vector&) = delete;
  device_vector& operator=(const device_vector&) = delete;
  device_vector(device_vector&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }

  void resize(size_t n) {
    if (n == n_) return;
    T* np = nullptr;
    if (n) check_hip(hipMalloc(&np, n * sizeof(T)), "hipMalloc");
    if (p_ && n_ && n) check_hip(hipMemcpy(np, p_, std::min(n, n_) * sizeof(T), hipMemcpyDeviceToDevice), "hipMemcpy");
    release();
    p_ = np;
    n_ = n;
  }
  void assign(const T* host, size_t n) {
    resize(n);
    if (n) check_hip(hipMemcpy(p_, host, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
  }
  std::vector<T> to_host() const {
    std::vector<T> h(n_);
    if (n_) check_hip(hipMemcpy(h.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
    return h;
  }
  T* data() { return p_; }
  const T* data() const { return p_; }
  size_t size() const { return n_; }

 private:
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* p_ = nullptr;
  size_t n_ = 0;
};

template <typename T>
struct pinned_allocator {
  using value_type = T;
  pinned_allocator() = default;
  template <typename U>
  pinned_allocator(const pinned_allocator<U>&) {}
  T* allocate(size_t n) {
    void* p = nullptr;
    check_hip(hipHostMalloc(&p, n * sizeof(T)), "hipHostMalloc");
    return static_cast<T*>(p);
  }
  void deallocate(T* p, size_t) { (void)hipHostFree(p); }
  template <typename U>
  bool operator==(const pinned_allocator<U>&) const { return true; }
  template <typename U>
  bool operator!=(const pinned_allocator<U>&) const { return false; }
};

template <typename T>
using pinned_vector = std::vector<T, pinned_allocator<T>>;

}  // namespace rtspatial
#endif
