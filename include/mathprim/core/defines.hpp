// mathprim/core/defines.hpp -- the part of mathprim's vocabulary that ssim's header-only API exposes, re-provided
// natively for the HIP build: index type, device tags, function qualifiers.
//
// The reference's mathprim (github.com/Adversarr/mathprim) is not vendored with ssim; its names are kept here so
// that code written against ssim::fem::basic_time_step<..., device::cuda, ...> ports by changing the device and
// BLAS template arguments.  Only the subset that ssim's hot path uses is provided.
#pragma once

#include <chrono>  // the reference drivers time steps with std::chrono through mathprim's headers
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MATHPRIM_HOST __host__
#define MATHPRIM_DEVICE __device__
#define MATHPRIM_GENERAL __host__ __device__
#else  // a host-only translation unit (g++): views and buffers of host memory still work
#define MATHPRIM_HOST
#define MATHPRIM_DEVICE
#define MATHPRIM_GENERAL
#endif
#define MATHPRIM_FORCE_INLINE inline
#define MATHPRIM_NOINLINE
#define MATHPRIM_PRIMFUNC MATHPRIM_GENERAL MATHPRIM_FORCE_INLINE
#define MATHPRIM_CONSTEXPR constexpr
#define MATHPRIM_UNUSED(x) (void)(x)
#define MATHPRIM_ASSERT(cond) ((void)0)
#define MATHPRIM_INTERNAL_CHECK_THROW(cond, expt, msg) \
  do {                                                 \
    if (!(cond)) throw expt(msg);                      \
  } while (0)
#define MATHPRIM_INTERNAL_COPY(t, enable) \
  t(const t&) = enable;                   \
  t& operator=(const t&) = enable
#define MATHPRIM_INTERNAL_MOVE(t, enable) \
  t(t&&) noexcept = enable;               \
  t& operator=(t&&) noexcept = enable

namespace mathprim {

// 32-bit indices: the C ABI underneath addresses vertices, cells and CSR entries with int32 (every benchmark
// configuration has < 2^31 nonzeros).
using index_t = int;
using size_t = std::size_t;

// compile-time extent of a view dimension; keep_dim marks a runtime extent
template <index_t N>
struct holder {
  static constexpr index_t value = N;
};
constexpr index_t keep_dim = -1;

namespace device {
struct cpu {};  // host memory
struct hip {};  // memory of the current HIP device (hipMalloc)
}  // namespace device

}  // namespace mathprim
