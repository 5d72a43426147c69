// mathprim/blas/blas.hpp -- basic_blas<Impl, T, Device>: the BLAS-1 interface of ssim's Blas template argument
// (copy, axpy, dot, norm, asum, scal on views of one device; finite_elements/time_step.hpp:490, 509-510, 538-542).
#pragma once

#include <cmath>

#include "mathprim/core/buffer.hpp"

namespace mathprim::blas {

template <typename Derived, typename T, typename Device>
class basic_blas {
 public:
  using scalar_type = T;
  using device_type = Device;
  template <typename V>
  using vec = contiguous_view<V, shape_t<keep_dim>, Device>;

  // y ← x
  template <typename A, typename SA, typename B, typename SB>
  void copy(const contiguous_view<A, SA, Device>& y, const contiguous_view<B, SB, Device>& x) {
    mathprim::copy(y, x);
  }
  // y ← α·x + y
  template <typename A, typename SA, typename B, typename SB>
  void axpy(T alpha, const contiguous_view<A, SA, Device>& x, const contiguous_view<B, SB, Device>& y) {
    check(x.numel(), y.numel());
    derived().axpy_impl(alpha, vec<const T>(x.data(), shape_t<keep_dim>(x.numel())), vec<T>(y.data(), shape_t<keep_dim>(y.numel())));
  }
  // x ← α·x
  template <typename A, typename SA>
  void scal(T alpha, const contiguous_view<A, SA, Device>& x) {
    derived().scal_impl(alpha, vec<T>(x.data(), shape_t<keep_dim>(x.numel())));
  }
  template <typename A, typename SA, typename B, typename SB>
  T dot(const contiguous_view<A, SA, Device>& x, const contiguous_view<B, SB, Device>& y) {
    check(x.numel(), y.numel());
    return derived().dot_impl(vec<const T>(x.data(), shape_t<keep_dim>(x.numel())),
                              vec<const T>(y.data(), shape_t<keep_dim>(y.numel())));
  }
  template <typename A, typename SA>
  T norm(const contiguous_view<A, SA, Device>& x) {
    return std::sqrt(dot(x, x));
  }
  template <typename A, typename SA>
  T asum(const contiguous_view<A, SA, Device>& x) {
    return derived().asum_impl(vec<const T>(x.data(), shape_t<keep_dim>(x.numel())));
  }

 private:
  Derived& derived() noexcept { return static_cast<Derived&>(*this); }
  static void check(index_t a, index_t b) {
    if (a != b) throw std::invalid_argument("blas: vectors of different lengths");
  }
};

// host BLAS-1 (plain loops), the Blas argument of a device::cpu instantiation
template <typename T>
class cpu_blas : public basic_blas<cpu_blas<T>, T, device::cpu> {
 public:
  void axpy_impl(T a, contiguous_vector_view<const T, device::cpu> x, contiguous_vector_view<T, device::cpu> y) {
    for (index_t i = 0; i < x.numel(); ++i) y[i] += a * x[i];
  }
  void scal_impl(T a, contiguous_vector_view<T, device::cpu> x) {
    for (index_t i = 0; i < x.numel(); ++i) x[i] *= a;
  }
  T dot_impl(contiguous_vector_view<const T, device::cpu> x, contiguous_vector_view<const T, device::cpu> y) {
    T s = 0;
    for (index_t i = 0; i < x.numel(); ++i) s += x[i] * y[i];
    return s;
  }
  T asum_impl(contiguous_vector_view<const T, device::cpu> x) {
    T s = 0;
    for (index_t i = 0; i < x.numel(); ++i) s += std::abs(x[i]);
    return s;
  }
};

}  // namespace mathprim::blas
