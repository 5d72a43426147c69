// ssim/finite_elements/stepper/nonlinear_cg.hpp -- nonlinear CG steppers (reference:
// finite_elements/stepper/nonlinear_cg.hpp:9-97): Dai-Kou β with the L-BFGS line-search constants, plain or
// preconditioned by z = L·Lᵀ·g with a user factor L on the system matrix's CSR pattern (identity after reset).
#pragma once

#include "ssim/finite_elements/time_step.hpp"

namespace ssim::fem {

class time_step_solver_ncg : public library_stepper<time_step_solver_ncg, SSIM_STEPPER_NCG> {};

template <typename Scalar, typename Device, typename SparseBlas>
class time_step_solver_ncg_with_ext_prec
    : public library_stepper<time_step_solver_ncg_with_ext_prec<Scalar, Device, SparseBlas>,
                             SSIM_STEPPER_NCG_EXT_PREC> {
 public:
  // the factor L: nnz values in the CSR order of sysmatrix() (bindings/src/fem_cuda.cu:204)
  template <typename Step>
  void set_matrix(Step& step, const Scalar* host_values, long long nnz) {
    internal::check(ssim_set_precond_matrix(step.handle(), host_values, nnz));
  }
};

}  // namespace ssim::fem
