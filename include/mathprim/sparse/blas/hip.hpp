// mathprim/sparse/blas/hip.hpp -- sparse::blas::hip<T, csr>: the SparseBlas argument of the HIP build, in the
// place of the reference's sparse::blas::cusparse (tests/fem_cu/main.cu:32).  basic_time_step assembles and
// multiplies its system matrix inside libssim_hip.so (symmetric blocked ELL), so this type mainly selects the
// device; gemv(α, x, β, y) multiplies a CSR matrix held in device memory, for callers that bring their own.
#pragma once

#include <hip/hip_runtime.h>

#include "mathprim/core/devices/hip.hpp"
#include "mathprim/sparse/basic_sparse.hpp"

namespace mathprim::sparse::blas {

namespace internal {
template <typename T>
__global__ void __launch_bounds__(256) csr_gemv_kernel(index_t rows, const T* val, const index_t* outer,
                                                       const index_t* inner, T alpha, const T* x, T beta, T* y) {
  const index_t r = static_cast<index_t>(blockIdx.x) * 256 + static_cast<index_t>(threadIdx.x);
  if (r >= rows) return;
  T s = 0;
  for (index_t k = outer[r]; k < outer[r + 1]; ++k) s += val[k] * x[inner[k]];
  y[r] = alpha * s + (beta == T(0) ? T(0) : beta * y[r]);
}
}  // namespace internal

template <typename T, sparse_format Format>
class hip;

template <typename T>
class hip<T, sparse_format::csr> {
 public:
  using matrix_view = basic_sparse_view<const T, device::hip, sparse_format::csr>;
  hip() = default;
  explicit hip(matrix_view A) : A_(A) {}
  // y ← α·A·x + β·y
  template <typename SX, typename SY>
  void gemv(T alpha, const contiguous_view<const T, SX, device::hip>& x, T beta,
            const contiguous_view<T, SY, device::hip>& y) const {
    if (x.numel() != A_.cols() || y.numel() != A_.rows()) throw std::invalid_argument("gemv: shape mismatch");
    if (A_.rows() == 0) return;
    internal::csr_gemv_kernel<<<(A_.rows() + 255) / 256, 256>>>(A_.rows(), A_.values().data(), A_.outer_ptrs().data(),
                                                                A_.inner_indices().data(), alpha, x.data(), beta,
                                                                y.data());
    hip_check(hipGetLastError(), "gemv launch");
    hip_check(hipDeviceSynchronize(), "gemv");
  }

 private:
  matrix_view A_{};
};

}  // namespace mathprim::sparse::blas
