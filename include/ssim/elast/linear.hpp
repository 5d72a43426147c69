// ssim/elast/linear.hpp -- elast::linear<Scalar, Device, Ndim> (reference: elast/linear.hpp): as the ElastModel
// of basic_time_step it selects the linear-elasticity element kernels (SSIM_MATERIAL_LINEAR).
#pragma once

#include "ssim/elast/basic_elast.hpp"

namespace ssim::elast {

template <typename Scalar, typename Device, index_t Ndim>
class linear : public basic_elast_model<linear<Scalar, Device, Ndim>, Scalar, Device, Ndim> {
  using base = basic_elast_model<linear<Scalar, Device, Ndim>, Scalar, Device, Ndim>;

 public:
  static constexpr ssim_material_t material = SSIM_MATERIAL_LINEAR;
  linear() = default;
  SSIM_PRIMFUNC linear(Scalar lambda, Scalar mu) noexcept : base(lambda, mu) {}
  // elast/basic_elast.hpp:225
  Scalar approximated_diffusivity() const noexcept { return this->lambda_ + 2 * this->mu_; }
};

}  // namespace ssim::elast
