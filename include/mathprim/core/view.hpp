// mathprim/core/view.hpp -- non-owning, row-major contiguous views: the view vocabulary of ssim's API
// (contiguous_view<T, shape_t<...>, Device>, batched<View>, contiguous_vector_view<T, Device>).
//
// A view is a pointer plus its extents.  It is a trivially copyable value, so device views are passed by value
// into kernels and __device__ lambdas, as the reference passes them (bindings/src/fem_cuda.cu:119).  Element
// access is row-major: view(i, j) of a [n, 3] view is data[3·i + j]; view[i] of a rank-r view (r > 1) is the
// rank-(r−1) view of row i, and of a rank-1 view the element i.
#pragma once

#include "mathprim/core/defines.hpp"

namespace mathprim {

// Extents of a view: compile-time values where known (holder<N>), runtime ones (keep_dim) otherwise; every
// extent is also stored, so shape(i) is a plain load on host and device.
template <index_t... Dims>
struct shape_t {
  static constexpr index_t ndim = static_cast<index_t>(sizeof...(Dims));
  index_t dims_[ndim > 0 ? ndim : 1];

  MATHPRIM_PRIMFUNC shape_t() noexcept : dims_{(Dims == keep_dim ? 0 : Dims)...} {}
  template <typename... I, typename = std::enable_if_t<sizeof...(I) == sizeof...(Dims)>>
  MATHPRIM_PRIMFUNC explicit shape_t(I... d) noexcept : dims_{static_cast<index_t>(d)...} {}

  MATHPRIM_PRIMFUNC index_t operator()(index_t i) const noexcept { return dims_[i]; }
  MATHPRIM_PRIMFUNC index_t at(index_t i) const noexcept { return dims_[i]; }
  MATHPRIM_PRIMFUNC index_t numel() const noexcept {
    index_t n = 1;
    for (index_t i = 0; i < ndim; ++i) n *= dims_[i];
    return n;
  }
  template <index_t... D2>
  MATHPRIM_PRIMFUNC bool operator==(const shape_t<D2...>& o) const noexcept {
    if (sizeof...(D2) != sizeof...(Dims)) return false;
    for (index_t i = 0; i < ndim; ++i)
      if (dims_[i] != o.dims_[i]) return false;
    return true;
  }
  template <index_t... D2>
  MATHPRIM_PRIMFUNC bool operator!=(const shape_t<D2...>& o) const noexcept {
    return !(*this == o);
  }
};

namespace internal {
template <typename S>
struct shape_tail;
template <index_t D0, index_t... Ds>
struct shape_tail<shape_t<D0, Ds...>> {
  using type = shape_t<Ds...>;
};
template <index_t D, typename S>
struct shape_prepend;
template <index_t D, index_t... Ds>
struct shape_prepend<D, shape_t<Ds...>> {
  using type = shape_t<D, Ds...>;
};
template <typename T>
struct extent_of {
  static constexpr index_t value = keep_dim;
  static MATHPRIM_PRIMFUNC index_t get(T v) noexcept { return static_cast<index_t>(v); }
};
template <index_t N>
struct extent_of<holder<N>> {
  static constexpr index_t value = N;
  static MATHPRIM_PRIMFUNC index_t get(holder<N>) noexcept { return N; }
};
// an extent argument: an integer or holder<N>
template <typename T>
struct is_extent : std::is_integral<T> {};
template <index_t N>
struct is_extent<holder<N>> : std::true_type {};
}  // namespace internal

// shape of runtime extents, e.g. make_shape(n, 3)
template <typename... I>
MATHPRIM_PRIMFUNC auto make_shape(I... d) noexcept {
  return shape_t<internal::extent_of<I>::value...>(internal::extent_of<I>::get(d)...);
}

template <typename T, typename Shape, typename Device>
class contiguous_view {
 public:
  using value_type = T;
  using shape_type = Shape;
  using device_type = Device;
  static constexpr index_t ndim = Shape::ndim;

  MATHPRIM_PRIMFUNC contiguous_view() noexcept : data_(nullptr), shape_() {}
  MATHPRIM_PRIMFUNC contiguous_view(T* data, const Shape& shape) noexcept : data_(data), shape_(shape) {}
  // a view of T converts to a view of const T
  template <typename U, typename = std::enable_if_t<std::is_same_v<const U, T> && !std::is_same_v<U, T>>>
  MATHPRIM_PRIMFUNC contiguous_view(const contiguous_view<U, Shape, Device>& o) noexcept
      : data_(o.data()), shape_(o.shape()) {}

  MATHPRIM_PRIMFUNC T* data() const noexcept { return data_; }
  MATHPRIM_PRIMFUNC const Shape& shape() const noexcept { return shape_; }
  MATHPRIM_PRIMFUNC index_t shape(index_t i) const noexcept { return shape_.dims_[i]; }
  MATHPRIM_PRIMFUNC index_t size() const noexcept { return shape_.numel(); }
  MATHPRIM_PRIMFUNC index_t numel() const noexcept { return shape_.numel(); }
  MATHPRIM_PRIMFUNC index_t stride(index_t i) const noexcept {
    index_t s = 1;
    for (index_t k = i + 1; k < ndim; ++k) s *= shape_.dims_[k];
    return s;
  }
  MATHPRIM_PRIMFUNC explicit operator bool() const noexcept { return data_ != nullptr; }

  // element (i0, i1, ...) with one index per dimension
  template <typename... I, typename = std::enable_if_t<sizeof...(I) == static_cast<size_t>(Shape::ndim)>>
  MATHPRIM_PRIMFUNC T& operator()(I... idx) const noexcept {
    const index_t ii[] = {static_cast<index_t>(idx)...};
    size_t off = 0;
    for (index_t k = 0; k < ndim; ++k) off = off * static_cast<size_t>(shape_.dims_[k]) + static_cast<size_t>(ii[k]);
    return data_[off];
  }

  // row i: the element of a rank-1 view, the rank-(r−1) view otherwise
  MATHPRIM_PRIMFUNC decltype(auto) operator[](index_t i) const noexcept {
    if constexpr (ndim == 1) {
      return static_cast<T&>(data_[i]);
    } else {
      using tail = typename internal::shape_tail<Shape>::type;
      tail t;
      for (index_t k = 1; k < ndim; ++k) t.dims_[k - 1] = shape_.dims_[k];
      return contiguous_view<T, tail, Device>(data_ + static_cast<size_t>(i) * static_cast<size_t>(stride(0)), t);
    }
  }

  MATHPRIM_PRIMFUNC contiguous_view<T, shape_t<keep_dim>, Device> flatten() const noexcept {
    return {data_, shape_t<keep_dim>(numel())};
  }
  template <typename... I>
  MATHPRIM_PRIMFUNC auto reshape(I... d) const noexcept {
    auto s = make_shape(d...);
    return contiguous_view<T, decltype(s), Device>(data_, s);
  }
  MATHPRIM_PRIMFUNC contiguous_view<const T, Shape, Device> as_const() const noexcept { return {data_, shape_}; }

 private:
  T* data_;
  Shape shape_;
};

// View with one more leading runtime dimension: batched<vertex view [3]> is the [nv, 3] view of all vertices.
template <typename View>
using batched = contiguous_view<typename View::value_type,
                                typename internal::shape_prepend<keep_dim, typename View::shape_type>::type,
                                typename View::device_type>;

template <typename T, typename Device>
using contiguous_vector_view = contiguous_view<T, shape_t<keep_dim>, Device>;

// view(ptr, shape) on a device
template <typename Device = device::cpu, typename T, index_t... Dims>
MATHPRIM_PRIMFUNC contiguous_view<T, shape_t<Dims...>, Device> view(T* data, const shape_t<Dims...>& shape) noexcept {
  return {data, shape};
}

}  // namespace mathprim
