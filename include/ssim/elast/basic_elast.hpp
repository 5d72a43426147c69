// ssim/elast/basic_elast.hpp -- Lamé parameters from (E, ν) (reference: elast/basic_elast.hpp:11-20) and the
// common base of the material models: a model holds uniform (λ, μ).  In the HIP build the model's energy,
// stress and Hessian are evaluated by libssim_hip.so's element kernels; the C++ type selects which.
#pragma once

#include "ssim/defines.hpp"
#include "ssim_hip.h"

namespace ssim::elast {

template <typename Scalar>
SSIM_PRIMFUNC Scalar compute_mu(Scalar E, Scalar nu) noexcept {
  return E / (2 * (1 + nu));
}
template <typename Scalar>
SSIM_PRIMFUNC Scalar compute_lambda(Scalar E, Scalar nu) noexcept {
  return E * nu / ((1 + nu) * (1 - 2 * nu));
}

template <typename Derived, typename Scalar, typename Device, index_t Ndim>
class basic_elast_model {
 public:
  using scalar_type = Scalar;
  using device_type = Device;
  static constexpr index_t ndim = Ndim;
  basic_elast_model() = default;
  SSIM_PRIMFUNC basic_elast_model(Scalar lambda, Scalar mu) noexcept : lambda_(lambda), mu_(mu) {}
  SSIM_PRIMFUNC Scalar lambda() const noexcept { return lambda_; }
  SSIM_PRIMFUNC Scalar mu() const noexcept { return mu_; }

 protected:
  Scalar lambda_{0}, mu_{0};
};

}  // namespace ssim::elast
