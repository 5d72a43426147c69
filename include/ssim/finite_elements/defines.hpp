// ssim/finite_elements/defines.hpp -- the CSR matrix types of the time step (reference:
// finite_elements/defines.hpp:5-12).
#pragma once

#include "mathprim/sparse/basic_sparse.hpp"
#include "ssim/defines.hpp"

namespace ssim::fem {

template <typename Scalar, typename Device>
using sparse_matrix = mp::sparse::basic_sparse_matrix<Scalar, Device, mp::sparse::sparse_format::csr>;
template <typename Scalar, typename Device>
using sparse_view = mp::sparse::basic_sparse_view<Scalar, Device, mp::sparse::sparse_format::csr>;

}  // namespace ssim::fem
