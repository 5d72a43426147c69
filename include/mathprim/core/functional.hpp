// mathprim/core/functional.hpp -- small functors applied by vmap: affine_transform x ← A·x + b on rows of N
// components (the reference drivers stretch the unit box with it, tests/fem_cu/main.cu:34-38).
#pragma once

#include "mathprim/core/view.hpp"

namespace mathprim::functional {

template <typename Scalar, typename Device, index_t N>
struct affine_transform {
  Scalar lin_[N * N];  // A, row-major; the identity unless set
  Scalar bias_[N];     // b; zero unless set

  MATHPRIM_PRIMFUNC affine_transform() noexcept {
    for (index_t i = 0; i < N; ++i) {
      bias_[i] = Scalar(0);
      for (index_t j = 0; j < N; ++j) lin_[i * N + j] = i == j ? Scalar(1) : Scalar(0);
    }
  }
  template <typename T, typename S>
  MATHPRIM_PRIMFUNC void operator()(const contiguous_view<T, S, Device>& x) const noexcept {
    Scalar y[N];
    for (index_t i = 0; i < N; ++i) {
      y[i] = bias_[i];
      for (index_t j = 0; j < N; ++j) y[i] += lin_[i * N + j] * x[j];
    }
    for (index_t i = 0; i < N; ++i) x[i] = y[i];
  }
};

}  // namespace mathprim::functional
