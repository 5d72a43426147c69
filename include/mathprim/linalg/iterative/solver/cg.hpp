// mathprim/linalg/iterative/solver/cg.hpp -- the conjugate-gradient solver type that parameterises
// time_step_solver_backward_euler (tests/fem_cu/main.cu:54).  The HIP Newton stepper solves H d = −g with the
// Jacobi-PCG of libssim_hip.so (x0 = 0, stop when ‖r‖₂ ≤ rtol·‖b‖₂, FP64 dot products and norms); this type
// records the choice and its preconditioner.
#pragma once

#include "mathprim/linalg/iterative/precond/diagonal.hpp"

namespace mathprim::sparse::iterative {

template <typename Scalar, typename Device, typename SparseBlas, typename Blas,
          typename Preconditioner = none_preconditioner<Scalar, Device>>
struct cg {
  using scalar_type = Scalar;
  using device_type = Device;
  using preconditioner_type = Preconditioner;
  static constexpr bool jacobi = Preconditioner::is_jacobi;
};

}  // namespace mathprim::sparse::iterative
