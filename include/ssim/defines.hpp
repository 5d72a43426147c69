// ssim/defines.hpp -- namespace ssim with the mathprim names it re-exports (index_t, mp = mathprim) and the
// SSIM_* qualifiers, as the reference's ssim/defines.hpp provides them.
#pragma once

#include "mathprim/core/defines.hpp"

#define SSIM_FORCE_INLINE MATHPRIM_FORCE_INLINE
#define SSIM_NOINLINE MATHPRIM_NOINLINE
#define SSIM_HOST MATHPRIM_HOST
#define SSIM_DEVICE MATHPRIM_DEVICE
#define SSIM_GENERAL MATHPRIM_GENERAL
#define SSIM_PRIMFUNC MATHPRIM_PRIMFUNC
#define SSIM_ASSERT(cond) MATHPRIM_ASSERT(cond)
#define SSIM_CONSTEXPR MATHPRIM_CONSTEXPR
#define SSIM_UNUSED(x) MATHPRIM_UNUSED(x)
#define SSIM_INTERNAL_CHECK_THROW(cond, expt, msg) MATHPRIM_INTERNAL_CHECK_THROW(cond, expt, msg)

namespace ssim {
namespace mp = ::mathprim;
using size_t = mp::size_t;
using index_t = mp::index_t;
}  // namespace ssim
