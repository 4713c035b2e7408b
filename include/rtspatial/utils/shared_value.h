// A device-resident scalar with host set/get (reference
// include/rtspatial/utils/shared_value.h:11-58); a SharedValue<unsigned long
// long> is the `arg` of the counting handler.
#ifndef RTSPATIAL_AMD_UTILS_SHARED_VALUE_H
#define RTSPATIAL_AMD_UTILS_SHARED_VALUE_H

#include <hip/hip_runtime_api.h>

#include "rtspatial/utils/exception.h"

namespace rtspatial {

template <typename T>
class SharedValue {
 public:
  SharedValue() {
    check_hip(hipMalloc(&d_, sizeof(T)), "hipMalloc");
    check_hip(hipHostMalloc(&h_, sizeof(T)), "hipHostMalloc");
  }
  ~SharedValue() {
    if (d_) (void)hipFree(d_);
    if (h_) (void)hipHostFree(h_);
  }
  SharedValue(const SharedValue&) = delete;
  SharedValue& operator=(const SharedValue&) = delete;
  void set(const T& v) { set(nullptr, v); }
  void set(hipStream_t stream, const T& v) {
    *h_ = v;
    check_hip(hipMemcpyAsync(d_, h_, sizeof(T), hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
    check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }
  T get() const { return get(nullptr); }
  T get(hipStream_t stream) const {
    check_hip(hipMemcpyAsync(h_, d_, sizeof(T), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
    check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    return *h_;
  }
  T* data() { return d_; }

 private:
  T* d_ = nullptr;
  T* h_ = nullptr;
};

}  // namespace rtspatial
#endif
