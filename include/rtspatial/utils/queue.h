// Result queue (reference include/rtspatial/utils/queue.h:60-113) with a 64-bit
// counter. DeviceObject() returns the device copy of its rts_dev_queue, the
// `arg` of the collecting handler. size() may exceed capacity(): the counter
// counts every result, stored or not.
#ifndef RTSPATIAL_AMD_UTILS_QUEUE_H
#define RTSPATIAL_AMD_UTILS_QUEUE_H

#include <hip/hip_runtime_api.h>

#include <vector>

#include "rts_c.h"
#include "rtspatial/utils/exception.h"
#include "rtspatial/utils/memory.h"
#include "rtspatial/utils/shared_value.h"

namespace rtspatial {

template <typename T1, typename T2>
struct pair {
  T1 first;
  T2 second;
};

inline pair<uint32_t, uint32_t> make_pair(uint32_t a, uint32_t b) { return {a, b}; }

class Queue {
 public:
  using value_type = pair<uint32_t, uint32_t>;
  static_assert(sizeof(value_type) == sizeof(rts_pair), "(geom_id, query_id)");

  void Init(size_t capacity) {
    data_.resize(capacity ? capacity : 1);
    counter_.set(0);
    rts_dev_queue q;
    q.data = reinterpret_cast<rts_pair*>(data_.data());
    q.capacity = data_.size();
    q.counter = counter_.data();
    dev_.assign(&q, 1);
  }
  void Clear(hipStream_t stream = nullptr) { counter_.set(stream, 0); }
  size_t size(hipStream_t stream = nullptr) const {
    return const_cast<SharedValue<unsigned long long>&>(counter_).get(stream);
  }
  size_t capacity() const { return data_.size(); }
  bool overflowed(hipStream_t stream = nullptr) const { return size(stream) > capacity(); }
  value_type* data() { return data_.data(); }
  void* DeviceObject() { return dev_.data(); }
  // The stored results (at most capacity()) copied to the host.
  std::vector<value_type> ToHost(hipStream_t stream = nullptr) const {
    const size_t k = std::min(size(stream), capacity());
    std::vector<value_type> h(k);
    if (k) check_hip(hipMemcpy(h.data(), data_.data(), k * sizeof(value_type), hipMemcpyDeviceToHost), "hipMemcpy");
    return h;
  }

 private:
  device_vector<value_type> data_;
  SharedValue<unsigned long long> counter_;
  device_vector<rts_dev_queue> dev_;
};

}  // namespace rtspatial
#endif
