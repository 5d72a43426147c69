// mathprim/blas/hip.hpp -- blas::hip<T>: BLAS-1 on device::hip views (the Blas argument of the HIP build, in the
// place of the reference's blas::cublas<T>, tests/fem_cu/main.cu:33).  Each call launches one kernel on the null
// stream and waits; dot and asum sum one partial per workgroup on the host in a fixed order, so results are
// deterministic.  The time step's own hot-path reductions live in libssim_hip.so, not here.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "mathprim/blas/blas.hpp"
#include "mathprim/core/devices/hip.hpp"

namespace mathprim::blas {

namespace internal {
constexpr int kBlasBlock = 256;
template <typename T>
__global__ void __launch_bounds__(kBlasBlock) axpy_kernel(index_t n, T a, const T* x, T* y) {
  const index_t i = static_cast<index_t>(blockIdx.x) * kBlasBlock + static_cast<index_t>(threadIdx.x);
  if (i < n) y[i] += a * x[i];
}
template <typename T>
__global__ void __launch_bounds__(kBlasBlock) scal_kernel(index_t n, T a, T* x) {
  const index_t i = static_cast<index_t>(blockIdx.x) * kBlasBlock + static_cast<index_t>(threadIdx.x);
  if (i < n) x[i] *= a;
}
// partial[b] = Σ over workgroup b of x_i·y_i (mode 0) or |x_i| (mode 1), summed in a fixed tree order
template <typename T>
__global__ void __launch_bounds__(kBlasBlock) reduce_kernel(index_t n, const T* x, const T* y, int mode, T* partial) {
  __shared__ T s[kBlasBlock];
  const index_t i = static_cast<index_t>(blockIdx.x) * kBlasBlock + static_cast<index_t>(threadIdx.x);
  T v = 0;
  if (i < n) v = mode == 0 ? x[i] * y[i] : (x[i] < 0 ? -x[i] : x[i]);
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = kBlasBlock / 2; w > 0; w /= 2) {
    if (static_cast<int>(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}
}  // namespace internal

template <typename T>
class hip : public basic_blas<hip<T>, T, device::hip> {
 public:
  using vec = contiguous_vector_view<T, device::hip>;
  using cvec = contiguous_vector_view<const T, device::hip>;
  void axpy_impl(T a, cvec x, vec y) {
    const index_t n = x.numel();
    if (n == 0) return;
    internal::axpy_kernel<<<blocks(n), internal::kBlasBlock>>>(n, a, x.data(), y.data());
    finish();
  }
  void scal_impl(T a, vec x) {
    const index_t n = x.numel();
    if (n == 0) return;
    internal::scal_kernel<<<blocks(n), internal::kBlasBlock>>>(n, a, x.data());
    finish();
  }
  T dot_impl(cvec x, cvec y) { return reduce(x, y, 0); }
  T asum_impl(cvec x) { return reduce(x, x, 1); }

 private:
  static int blocks(index_t n) { return static_cast<int>((n + internal::kBlasBlock - 1) / internal::kBlasBlock); }
  static void finish() {
    hip_check(hipGetLastError(), "blas::hip launch");
    hip_check(hipDeviceSynchronize(), "blas::hip");
  }
  T reduce(cvec x, cvec y, int mode) {
    const index_t n = x.numel();
    if (n == 0) return T(0);
    const int nb = blocks(n);
    if (static_cast<int>(part_.size()) < nb) part_ = make_buffer<T, device::hip>(nb);
    internal::reduce_kernel<<<nb, internal::kBlasBlock>>>(n, x.data(), y.data(), mode, part_.data());
    finish();
    std::vector<T> h(nb);
    hip_check(hipMemcpy(h.data(), part_.data(), sizeof(T) * nb, hipMemcpyDeviceToHost), "hipMemcpy");
    T s = 0;
    for (T v : h) s += v;
    return s;
  }
  contiguous_buffer<T, shape_t<keep_dim>, device::hip> part_;
};

}  // namespace mathprim::blas
