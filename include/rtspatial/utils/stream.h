// RAII HIP stream (reference include/rtspatial/utils/stream.h:11-59).
#ifndef RTSPATIAL_AMD_UTILS_STREAM_H
#define RTSPATIAL_AMD_UTILS_STREAM_H

#include <hip/hip_runtime_api.h>

#include "rtspatial/utils/exception.h"

namespace rtspatial {

class Stream {
 public:
  Stream() { check_hip(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking), "hipStreamCreate"); }
  ~Stream() {
    if (s_) (void)hipStreamDestroy(s_);
  }
  Stream(const Stream&) = delete;
  Stream& operator=(const Stream&) = delete;
  hipStream_t hip_stream() const { return s_; }
  // Same accessor name as the reference's Stream::cuda_stream().
  hipStream_t cuda_stream() const { return s_; }
  void Sync() const { check_hip(hipStreamSynchronize(s_), "hipStreamSynchronize"); }

 private:
  hipStream_t s_ = nullptr;
};

}  // namespace rtspatial
#endif
