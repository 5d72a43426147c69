// mathprim/optim/basic_optim.hpp -- optim_result, the record a stepper keeps of its last solve
// (ssim/finite_elements/stepper/backward_euler.hpp:28-32, 40).
#pragma once

#include "mathprim/core/defines.hpp"

namespace mathprim::optim {

template <typename Scalar>
struct optim_result {
  Scalar value_{0};        // objective (incremental potential) at the result
  Scalar last_change_{0};  // |change of the objective| over the last accepted iteration
  Scalar grad_norm_{0};    // ‖gradient‖₂ at the result
  index_t iterations_{0};
  bool converged_{false};
};

}  // namespace mathprim::optim
