// ssim/finite_elements/stepper/backward_euler.hpp -- the Newton and L-BFGS steppers of the backward-Euler step
// (reference: finite_elements/stepper/backward_euler.hpp:9-67).
//
// time_step_solver_backward_euler<SparseSolver>: Newton with backtracking Armijo line search (c = 0.3), stopping
// at ‖g‖₂ < threshold·mass; the linear solve is the library's PCG, preconditioned as SparseSolver says
// (mp::sparse::iterative::cg<..., diagonal_preconditioner<...>> → Jacobi, cg<...> without one → plain CG).
// time_step_solver_lbfgs: L-BFGS, m = 8, H0 = the diagonal of the rest Hessian.
// Both keep `result` (mathprim optim_result) of their last solve.
#pragma once

#include "mathprim/linalg/iterative/solver/cg.hpp"
#include "ssim/finite_elements/time_step.hpp"

namespace ssim::fem {

template <typename SparseSolver>
class time_step_solver_backward_euler
    : public library_stepper<time_step_solver_backward_euler<SparseSolver>, SSIM_STEPPER_NEWTON> {
 public:
  int precond() const noexcept { return SparseSolver::jacobi ? 0 : 1; }
};

class time_step_solver_lbfgs : public library_stepper<time_step_solver_lbfgs, SSIM_STEPPER_LBFGS> {};

}  // namespace ssim::fem
