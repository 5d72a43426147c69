// mathprim/core/buffer.hpp -- owning contiguous buffers (make_buffer, to_buffer_t, fill_bytes) and copy()
// between views, on the host and (with mathprim/core/devices/hip.hpp) on the HIP device.
#pragma once

#include <cstdlib>
#include <cstring>
#include <new>

#include "mathprim/core/view.hpp"

namespace mathprim {

// Allocation and byte fill of one device; the HIP specialisation lives in mathprim/core/devices/hip.hpp.
template <typename Device>
struct device_traits;

template <>
struct device_traits<device::cpu> {
  static void* alloc(size_t bytes) {
    void* p = std::aligned_alloc(64, (bytes + 63) / 64 * 64);
    if (!p && bytes) throw std::bad_alloc();
    return p;
  }
  static void free(void* p) noexcept { std::free(p); }
  static void memset(void* p, int v, size_t bytes) { std::memset(p, v, bytes); }
};

// Byte copies between two devices; cpu ↔ cpu here, every pair with a HIP side in devices/hip.hpp.
template <typename DstDevice, typename SrcDevice>
struct copy_traits;

template <>
struct copy_traits<device::cpu, device::cpu> {
  static void copy(void* dst, const void* src, size_t bytes) { std::memmove(dst, src, bytes); }
};

template <typename T, typename Shape, typename Device>
class contiguous_buffer {
 public:
  using value_type = T;
  using shape_type = Shape;
  using device_type = Device;
  using view_type = contiguous_view<T, Shape, Device>;
  using const_view_type = contiguous_view<const T, Shape, Device>;

  contiguous_buffer() = default;
  explicit contiguous_buffer(const Shape& shape) : shape_(shape) {
    const size_t bytes = sizeof(T) * static_cast<size_t>(shape.numel());
    data_ = bytes ? static_cast<T*>(device_traits<Device>::alloc(bytes)) : nullptr;
  }
  ~contiguous_buffer() { reset(); }
  contiguous_buffer(const contiguous_buffer&) = delete;
  contiguous_buffer& operator=(const contiguous_buffer&) = delete;
  contiguous_buffer(contiguous_buffer&& o) noexcept : data_(o.data_), shape_(o.shape_) { o.data_ = nullptr; }
  contiguous_buffer& operator=(contiguous_buffer&& o) noexcept {
    if (this != &o) {
      reset();
      data_ = o.data_;
      shape_ = o.shape_;
      o.data_ = nullptr;
    }
    return *this;
  }

  view_type view() noexcept { return {data_, shape_}; }
  const_view_type view() const noexcept { return {data_, shape_}; }
  const_view_type const_view() const noexcept { return {data_, shape_}; }
  const Shape& shape() const noexcept { return shape_; }
  index_t shape(index_t i) const noexcept { return shape_(i); }
  index_t size() const noexcept { return shape_.numel(); }
  T* data() noexcept { return data_; }
  const T* data() const noexcept { return data_; }
  explicit operator bool() const noexcept { return data_ != nullptr; }
  void fill_bytes(int value) {
    if (data_) device_traits<Device>::memset(data_, value, sizeof(T) * static_cast<size_t>(size()));
  }

 private:
  void reset() noexcept {
    if (data_) device_traits<Device>::free(data_);
    data_ = nullptr;
  }
  T* data_ = nullptr;
  Shape shape_{};
};

template <typename View>
using to_buffer_t = contiguous_buffer<std::remove_const_t<typename View::value_type>, typename View::shape_type,
                                      typename View::device_type>;

// make_buffer<T, Device>(extents...) with runtime extents (integers) or compile-time ones (holder<N>), or from a
// shape object: make_buffer<double>(n, holder<3>{}) is an [n, 3] host buffer.
template <typename T, typename Device = device::cpu, index_t... Dims>
contiguous_buffer<T, shape_t<Dims...>, Device> make_buffer(const shape_t<Dims...>& shape) {
  return contiguous_buffer<T, shape_t<Dims...>, Device>(shape);
}
template <typename T, typename Device = device::cpu, typename... I,
          typename = std::enable_if_t<(sizeof...(I) > 0) && (internal::is_extent<I>::value && ...)>>
auto make_buffer(I... extents) {
  return make_buffer<T, Device>(make_shape(extents...));
}

// copy(dst, src): element-wise copy of two views with the same number of elements, across devices
template <typename T1, typename S1, typename D1, typename T2, typename S2, typename D2>
void copy(const contiguous_view<T1, S1, D1>& dst, const contiguous_view<T2, S2, D2>& src) {
  static_assert(std::is_same_v<std::remove_const_t<T1>, std::remove_const_t<T2>>, "copy between different types");
  static_assert(!std::is_const_v<T1>, "copy into a const view");
  if (dst.numel() != src.numel()) throw std::invalid_argument("copy: views hold different numbers of elements");
  if (dst.numel() == 0) return;
  copy_traits<D1, D2>::copy(dst.data(), src.data(), sizeof(T1) * static_cast<size_t>(dst.numel()));
}

// zeros(view): fill with zero bytes
template <typename T, typename S, typename D>
void zeros(const contiguous_view<T, S, D>& v) {
  if (v.numel()) device_traits<D>::memset(v.data(), 0, sizeof(T) * static_cast<size_t>(v.numel()));
}

}  // namespace mathprim
