// ssim/finite_elements/boundary.hpp -- node_boundary_type, the per-dof flag of basic_time_step::dof_type()
// (reference: finite_elements/boundary.hpp:8-11).  One byte per dof, 0 = general, 1 = dirichlet: the layout of
// libssim_hip.so's device flags, so a dof_type() view writes them directly.
#pragma once

#include "ssim/defines.hpp"

namespace ssim::fem {

enum class node_boundary_type : char {
  general,
  dirichlet,
};

}  // namespace ssim::fem
