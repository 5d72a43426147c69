// ssim/ssim.hpp -- everything of the HIP build's header-only API: basic_time_step on device::hip, its steppers,
// material models and mesh generators, and the mathprim subset they use.
#pragma once

#include "mathprim/blas/hip.hpp"
#include "mathprim/core/functional.hpp"
#include "mathprim/linalg/iterative/precond/diagonal.hpp"
#include "mathprim/linalg/iterative/solver/cg.hpp"
#include "mathprim/parallel/hip.hpp"
#include "mathprim/sparse/blas/hip.hpp"
#include "ssim/elast/linear.hpp"
#include "ssim/elast/stable_neohookean.hpp"
#include "ssim/finite_elements/stepper/backward_euler.hpp"
#include "ssim/finite_elements/stepper/nonlinear_cg.hpp"
#include "ssim/finite_elements/stepper/projective_dynamics.hpp"
#include "ssim/finite_elements/time_step.hpp"
#include "ssim/mesh/common_shapes.hpp"
