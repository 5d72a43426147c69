// ssim/finite_elements/stepper/projective_dynamics.hpp -- L-BFGS steppers whose two-loop middle step solves with
// a sparse matrix (reference: finite_elements/stepper/projective_dynamics.hpp:9-150):
// time_step_solver_pd<LinSolver> with the projective-dynamics matrix M + dt²κK built at reset, and
// time_step_solver_h0<LinSolver> with the SPD Hessian at the start of each step.  The reference factorises with
// LinSolver (CHOLMOD / Eigen LDLᵀ); the library solves by Jacobi-PCG to a relative residual of 1e-12, so
// LinSolver only names the reference's choice.
#pragma once

#include "ssim/finite_elements/time_step.hpp"

namespace ssim::fem {

template <typename LinSolver>
class time_step_solver_pd : public library_stepper<time_step_solver_pd<LinSolver>, SSIM_STEPPER_LBFGS_PD> {};

template <typename LinSolver>
class time_step_solver_h0 : public library_stepper<time_step_solver_h0<LinSolver>, SSIM_STEPPER_LBFGS_H0> {};

}  // namespace ssim::fem
