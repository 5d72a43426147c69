// mathprim/linalg/iterative/precond/diagonal.hpp -- the Jacobi (diagonal) preconditioner type of the reference's
// Newton drivers (tests/fem_cu/main.cu:52-53).  With the HIP build the preconditioner is applied inside
// libssim_hip.so's PCG (inverse diagonal written by the Hessian assembly); the type selects it.
#pragma once

#include "mathprim/sparse/basic_sparse.hpp"

namespace mathprim::sparse::iterative {

template <typename Scalar, typename Device, sparse_format Format, typename Blas>
struct diagonal_preconditioner {
  using scalar_type = Scalar;
  using device_type = Device;
  static constexpr bool is_jacobi = true;
};

// no preconditioning: not provided by the HIP build (its linear solve is Jacobi-PCG), kept as a name so that
// selecting it fails at compile time with a message instead of silently switching algorithm
template <typename Scalar, typename Device>
struct none_preconditioner {
  using scalar_type = Scalar;
  using device_type = Device;
  static constexpr bool is_jacobi = false;
};

}  // namespace mathprim::sparse::iterative
