// noalloc.hip -- ssim_pcg_solve_noalloc: the Jacobi-PCG solve of ssim_pcg_solve on caller-owned device vectors,
// with every work buffer in a static device array instead of memory the library allocates.
//
// For callers that cannot call hipMalloc (graph-capture setups, serving loops, harnesses that forbid it).  The
// work of one solve -- r, z, p, q (4·n doubles), the partial sums of the two PCG kernels and the device-side
// PCG scalars and iteration counter -- lives in `scratch`, a file-scope __device__ array of kScratchBytes
// (64 MiB, up to about 2.0M dofs), which each GPU the library runs on holds until the process ends.  Its
// per-device address is looked up once per device at ssim_create (the library's set-up call) with that device
// current.  Every call on a GPU shares that GPU's array; the kernels write every scratch word before reading
// it, so nothing carries over between calls, and calls that overlap on one GPU race.
#include <mutex>

#include "context.h"

using namespace ssim_hip;

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

void* g_scratch[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
std::once_flag g_scratch_once[kMaxDevices];

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// byte offsets of the work buffers of a solve with n dofs and pmax partials per partial array
struct ScratchLayout {
  size_t sc, it, pa, pb0, pb1, r, z, p, q, total;
  ScratchLayout(size_t n, size_t pmax) {
    sc = 0;
    it = align256(sc + sizeof(PcgScalars));
    pa = align256(it + sizeof(PcgIter));
    pb0 = align256(pa + sizeof(double) * pmax);
    pb1 = align256(pb0 + 2 * sizeof(double) * pmax);
    r = align256(pb1 + 2 * sizeof(double) * pmax);
    z = align256(r + sizeof(double) * n);
    p = align256(z + sizeof(double) * n);
    q = align256(p + sizeof(double) * n);
    total = align256(q + sizeof(double) * n);
  }
};

}  // namespace

// Looks up the current device's address of `scratch`, once per device.  A failure is recorded and reported by
// the entry point; contexts without it work as before.
void ssim_hip::noalloc_scratch_init() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return;
  std::call_once(g_scratch_once[dev], [dev] {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[dev] = e;
    g_scratch[dev] = e == hipSuccess ? p : nullptr;
  });
}

extern "C" {

ssim_status_t ssim_pcg_solve_noalloc(ssim_ctx* ctx, const double* b, double* x, double rtol, int maxit,
                                     int* iters_out) {
  if (!ctx) {
    g_err = "ctx is null";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  return run([&] {
    require_reset(ctx);
    if (!b || !x) fail(SSIM_ERR_INVALID_ARGUMENT, "null vector");
    if (!(rtol >= 0) || maxit < 0) fail(SSIM_ERR_INVALID_ARGUMENT, "rtol must be ≥ 0 and maxit ≥ 0");
    const ScratchLayout L((size_t)ctx->n, ctx->pmax);
    if (L.total > kScratchBytes)
      fail(SSIM_ERR_INVALID_ARGUMENT, "the solve needs " + std::to_string(L.total) + " bytes of scratch; the no-alloc " +
                                          "entry point holds " + std::to_string(kScratchBytes) + " (about " +
                                          std::to_string(kScratchBytes / 32 / 1000) + "k dofs)");
    const int dev = ctx->device;
    if (dev < 0 || dev >= kMaxDevices || !g_scratch[dev])
      fail(SSIM_ERR_HIP, std::string("no scratch address for this device (hipGetSymbolAddress: ") +
                             hipGetErrorString(dev >= 0 && dev < kMaxDevices ? g_scratch_err[dev] : hipErrorInvalidDevice) +
                             ")");
    unsigned char* base = static_cast<unsigned char*>(g_scratch[dev]);
    const PcgWork w{reinterpret_cast<double*>(base + L.r),   reinterpret_cast<double*>(base + L.z),
                    reinterpret_cast<double*>(base + L.p),   reinterpret_cast<double*>(base + L.q),
                    reinterpret_cast<double*>(base + L.pa),  reinterpret_cast<double*>(base + L.pb0),
                    reinterpret_cast<double*>(base + L.pb1), reinterpret_cast<PcgScalars*>(base + L.sc),
                    reinterpret_cast<PcgIter*>(base + L.it)};
    const int it = pcg_solve(ctx, b, 1.0, x, rtol, maxit, nullptr, nullptr, &w);
    if (iters_out) *iters_out = it;
  });
}

}  // extern "C"
