// Non-owning (pointer, size) view (reference include/rtspatial/utils/array_view.h:10-69).
// May view device memory or pinned host memory.
#ifndef RTSPATIAL_AMD_UTILS_ARRAY_VIEW_H
#define RTSPATIAL_AMD_UTILS_ARRAY_VIEW_H

#include <cstddef>
#include <utility>

namespace rtspatial {

template <typename T>
class ArrayView {
 public:
  ArrayView() = default;
  ArrayView(T* data, size_t size) : data_(data), size_(size) {}
  template <typename V, typename = decltype(std::declval<V&>().data())>
  ArrayView(V& v) : data_(const_cast<T*>(v.data())), size_(v.size()) {}
  T* data() { return data_; }
  const T* data() const { return data_; }
  size_t size() const { return size_; }
  bool empty() const { return size_ == 0; }
  T* begin() { return data_; }
  T* end() { return data_ + size_; }

 private:
  T* data_ = nullptr;
  size_t size_ = 0;
};

}  // namespace rtspatial
#endif
