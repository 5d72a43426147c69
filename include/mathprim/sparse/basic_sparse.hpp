// mathprim/sparse/basic_sparse.hpp -- CSR matrices and views (the sys_matrix / sparse_view types of
// ssim/finite_elements/defines.hpp): values, outer_ptrs, inner_indices, rows, cols, nnz, and visit(f), a task
// that calls f(row, col, value&) for every stored entry when a parfor runs it (tests/fem_cu/main.cu:68-72).
#pragma once

#include <vector>

#include "mathprim/core/buffer.hpp"

namespace mathprim::sparse {

enum class sparse_format { csr, csc, coo };
enum class sparse_property { general, symmetric };

template <typename T, typename Device, sparse_format Format>
class basic_sparse_view;

// the task visit() returns: par.run(task) calls f(row, col, value) over the stored entries in row order
template <typename T, typename F>
struct csr_visit_task {
  T* values;
  const index_t* outer;
  const index_t* inner;
  index_t rows;
  F f;
  template <typename Par>
  void run_impl(const Par&) {
    for (index_t r = 0; r < rows; ++r)
      for (index_t k = outer[r]; k < outer[r + 1]; ++k) f(r, inner[k], values[k]);
  }
};

template <typename T, typename Device>
class basic_sparse_view<T, Device, sparse_format::csr> {
 public:
  using value_type = T;
  using device_type = Device;
  using index_view = contiguous_vector_view<const index_t, Device>;
  static constexpr sparse_format format = sparse_format::csr;

  basic_sparse_view() = default;
  basic_sparse_view(T* values, const index_t* outer_ptrs, const index_t* inner_indices, index_t rows, index_t cols,
                    index_t nnz, sparse_property property = sparse_property::general) noexcept
      : values_(values), outer_(outer_ptrs), inner_(inner_indices), rows_(rows), cols_(cols), nnz_(nnz),
        property_(property) {}

  contiguous_vector_view<T, Device> values() const noexcept { return {values_, shape_t<keep_dim>(nnz_)}; }
  index_view outer_ptrs() const noexcept { return {outer_, shape_t<keep_dim>(rows_ + 1)}; }
  index_view inner_indices() const noexcept { return {inner_, shape_t<keep_dim>(nnz_)}; }
  index_t rows() const noexcept { return rows_; }
  index_t cols() const noexcept { return cols_; }
  index_t nnz() const noexcept { return nnz_; }
  sparse_property property() const noexcept { return property_; }
  basic_sparse_view<const T, Device, sparse_format::csr> as_const() const noexcept {
    return {values_, outer_, inner_, rows_, cols_, nnz_, property_};
  }
  // visiting runs on the host: a view of device::cpu memory only
  template <typename F>
  csr_visit_task<T, F> visit(F f) const {
    static_assert(std::is_same_v<Device, device::cpu>, "visit() reads the entries on the host");
    return {values_, outer_, inner_, rows_, f};
  }

 private:
  T* values_ = nullptr;
  const index_t* outer_ = nullptr;
  const index_t* inner_ = nullptr;
  index_t rows_ = 0, cols_ = 0, nnz_ = 0;
  sparse_property property_ = sparse_property::general;
};

// Owning CSR matrix (host memory).
template <typename T, typename Device, sparse_format Format>
class basic_sparse_matrix;

template <typename T>
class basic_sparse_matrix<T, device::cpu, sparse_format::csr> {
 public:
  basic_sparse_matrix() = default;
  basic_sparse_matrix(index_t rows, index_t cols, index_t nnz, sparse_property property = sparse_property::general)
      : values_(static_cast<size_t>(nnz)), outer_(static_cast<size_t>(rows) + 1), inner_(static_cast<size_t>(nnz)),
        rows_(rows), cols_(cols), property_(property) {}
  basic_sparse_view<T, device::cpu, sparse_format::csr> view() noexcept {
    return {values_.data(), outer_.data(), inner_.data(), rows_, cols_, nnz(), property_};
  }
  basic_sparse_view<const T, device::cpu, sparse_format::csr> const_view() const noexcept {
    return {values_.data(), outer_.data(), inner_.data(), rows_, cols_, nnz(), property_};
  }
  basic_sparse_view<const T, device::cpu, sparse_format::csr> view() const noexcept { return const_view(); }
  index_t rows() const noexcept { return rows_; }
  index_t cols() const noexcept { return cols_; }
  index_t nnz() const noexcept { return static_cast<index_t>(values_.size()); }
  T* values_data() noexcept { return values_.data(); }
  index_t* outer_data() noexcept { return outer_.data(); }
  index_t* inner_data() noexcept { return inner_.data(); }

 private:
  std::vector<T> values_;
  std::vector<index_t> outer_, inner_;
  index_t rows_ = 0, cols_ = 0;
  sparse_property property_ = sparse_property::general;
};

}  // namespace mathprim::sparse
