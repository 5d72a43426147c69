// mathprim/parallel/hip.hpp -- par::hip, the ParImpl of the HIP build: run(n, f) launches one thread per index
// on the calling thread's HIP stream (default: the null stream) and waits for it, as the reference's par::cuda
// does for its device lambdas (tests/fem_cu/main.cu:41-45).  f must be callable on the device (a __device__ or
// __host__ __device__ lambda or functor whose captures are device views and values).
#pragma once

#include <hip/hip_runtime.h>

#include "mathprim/core/devices/hip.hpp"
#include "mathprim/parallel/parallel.hpp"

namespace mathprim::par {

namespace internal {
template <typename F>
__global__ void __launch_bounds__(256) run_kernel(index_t n, F f) {
  const index_t i = static_cast<index_t>(blockIdx.x) * 256 + static_cast<index_t>(threadIdx.x);
  if (i < n) f(i);
}
template <typename F, typename View0, typename... Views>
__global__ void __launch_bounds__(256) vmap_kernel(index_t n, F f, View0 v0, Views... vs) {
  const index_t i = static_cast<index_t>(blockIdx.x) * 256 + static_cast<index_t>(threadIdx.x);
  if (i < n) f(v0[i], vs[i]...);
}
}  // namespace internal

class hip : public parfor<hip> {
 public:
  hip() = default;
  explicit hip(hipStream_t stream) : stream_(stream) {}
  hipStream_t stream() const noexcept { return stream_; }

  template <typename F>
  void run_impl(index_t n, F f) const {
    if (n <= 0) return;
    internal::run_kernel<<<(n + 255) / 256, 256, 0, stream_>>>(n, f);
    finish();
  }
  template <typename F, typename View0, typename... Views>
  void vmap_impl(F f, const View0& v0, const Views&... vs) const {
    const index_t n = v0.shape(0);
    if (n <= 0) return;
    internal::vmap_kernel<<<(n + 255) / 256, 256, 0, stream_>>>(n, f, v0, vs...);
    finish();
  }

 private:
  void finish() const {
    hip_check(hipGetLastError(), "par::hip launch");
    hip_check(hipStreamSynchronize(stream_), "par::hip");
  }
  hipStream_t stream_ = nullptr;
};

}  // namespace mathprim::par
