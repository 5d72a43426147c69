// mathprim/core/devices/hip.hpp -- device::hip: buffers in the memory of the current HIP device (hipMalloc) and
// copies between host and device memory (hipMemcpy with the direction inferred from the pointers).
#pragma once

#include <hip/hip_runtime_api.h>

#include "mathprim/core/buffer.hpp"

namespace mathprim {

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <>
struct device_traits<device::hip> {
  static void* alloc(size_t bytes) {
    void* p = nullptr;
    hip_check(hipMalloc(&p, bytes), "hipMalloc");
    return p;
  }
  static void free(void* p) noexcept { (void)hipFree(p); }
  static void memset(void* p, int v, size_t bytes) { hip_check(hipMemset(p, v, bytes), "hipMemset"); }
};

namespace internal {
inline void hip_copy(void* dst, const void* src, size_t bytes) {
  hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDefault), "hipMemcpy");
}
}  // namespace internal

template <>
struct copy_traits<device::hip, device::hip> {
  static void copy(void* d, const void* s, size_t n) { internal::hip_copy(d, s, n); }
};
template <>
struct copy_traits<device::hip, device::cpu> {
  static void copy(void* d, const void* s, size_t n) { internal::hip_copy(d, s, n); }
};
template <>
struct copy_traits<device::cpu, device::hip> {
  static void copy(void* d, const void* s, size_t n) { internal::hip_copy(d, s, n); }
};

}  // namespace mathprim
