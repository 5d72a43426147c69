// ssim/elast/stable_neohookean.hpp -- elast::stable_neohookean<Scalar, Device, Ndim> (reference:
// elast/stable_neohookean.hpp): as the ElastModel of basic_time_step it selects the stable Neo-Hookean element
// kernels (SSIM_MATERIAL_STABLE_NEOHOOKEAN).
#pragma once

#include "ssim/elast/basic_elast.hpp"

namespace ssim::elast {

template <typename Scalar, typename Device, index_t Ndim>
class stable_neohookean : public basic_elast_model<stable_neohookean<Scalar, Device, Ndim>, Scalar, Device, Ndim> {
  using base = basic_elast_model<stable_neohookean<Scalar, Device, Ndim>, Scalar, Device, Ndim>;

 public:
  static constexpr ssim_material_t material = SSIM_MATERIAL_STABLE_NEOHOOKEAN;
  stable_neohookean() = default;
  SSIM_PRIMFUNC stable_neohookean(Scalar lambda, Scalar mu) noexcept : base(lambda, mu) {}
  // diffusivity of the projective-dynamics matrix (elast/stable_neohookean.hpp:109-114)
  Scalar approximated_diffusivity() const noexcept {
    return (this->lambda_ + Scalar(5.0 / 6.0) * this->mu_) + 2 * (Scalar(4.0 / 3.0) * this->mu_);
  }
};

}  // namespace ssim::elast
