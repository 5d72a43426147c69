// ssim/mesh/basic_mesh.hpp -- basic_unstructured<Scalar, Device, PhysicalDim, TopologyDim>: a mesh that owns its
// vertices [nv, PhysicalDim] and cells [nt, TopologyDim] on one device, with its non-owning view type; tet_mesh
// is the tetrahedral case (reference: ssim/include/ssim/mesh/basic_mesh.hpp:20-129).
#pragma once

#include "mathprim/core/buffer.hpp"
#include "ssim/defines.hpp"

namespace ssim {

template <typename Scalar, typename Device, index_t PhysicalDim, index_t TopologyDim>
class basic_unstructured_view {
 public:
  using scalar_type = Scalar;
  using device_type = Device;
  static constexpr index_t physical_dim = PhysicalDim;
  static constexpr index_t topology_dim = TopologyDim;
  static constexpr bool is_const = std::is_const_v<Scalar>;
  using maybe_const_index = std::conditional_t<is_const, const index_t, index_t>;
  using vertex_type = mp::contiguous_view<Scalar, mp::shape_t<physical_dim>, Device>;
  using cell_type = mp::contiguous_view<maybe_const_index, mp::shape_t<topology_dim>, Device>;
  using batched_vertex = mp::batched<vertex_type>;
  using batched_cell = mp::batched<cell_type>;

  basic_unstructured_view() = default;
  SSIM_PRIMFUNC basic_unstructured_view(const batched_vertex& vertices, const batched_cell& cells) noexcept
      : vertices_(vertices), cells_(cells) {}
  SSIM_PRIMFUNC const batched_vertex& vertices() const noexcept { return vertices_; }
  SSIM_PRIMFUNC const batched_cell& cells() const noexcept { return cells_; }
  SSIM_PRIMFUNC index_t num_vertices() const noexcept { return vertices_.shape(0); }
  SSIM_PRIMFUNC index_t num_cells() const noexcept { return cells_.shape(0); }
  SSIM_PRIMFUNC basic_unstructured_view<std::add_const_t<Scalar>, Device, PhysicalDim, TopologyDim> as_const()
      const noexcept {
    return {vertices_.as_const(), cells_.as_const()};
  }

 private:
  batched_vertex vertices_;
  batched_cell cells_;
};

template <typename Scalar, typename Device, index_t PhysicalDim, index_t TopologyDim>
class basic_unstructured {
  static_assert(!std::is_const_v<Scalar>, "Scalar type must be non-const.");

 public:
  using scalar_type = Scalar;
  using device_type = Device;
  static constexpr index_t physical_dim = PhysicalDim;
  static constexpr index_t topology_dim = TopologyDim;
  using view_type = basic_unstructured_view<Scalar, Device, PhysicalDim, TopologyDim>;
  using const_view_type = basic_unstructured_view<const Scalar, Device, PhysicalDim, TopologyDim>;
  using vertex_type = typename view_type::vertex_type;
  using cell_type = typename view_type::cell_type;
  using batched_vertex = typename view_type::batched_vertex;
  using batched_cell = typename view_type::batched_cell;
  using batched_vertex_buffer = mp::to_buffer_t<batched_vertex>;
  using batched_cell_buffer = mp::to_buffer_t<batched_cell>;

  basic_unstructured() = default;
  basic_unstructured(const basic_unstructured&) = delete;  // meshes move (the time step takes ownership)
  basic_unstructured& operator=(const basic_unstructured&) = delete;
  basic_unstructured(basic_unstructured&&) noexcept = default;
  basic_unstructured& operator=(basic_unstructured&&) noexcept = default;

  basic_unstructured(index_t num_vertices, index_t num_cells)
      : vertices_(mp::make_buffer<Scalar, Device>(num_vertices, mp::holder<physical_dim>{})),
        cells_(mp::make_buffer<index_t, Device>(num_cells, mp::holder<topology_dim>{})) {}

  view_type view() noexcept { return {vertices_.view(), cells_.view()}; }
  const_view_type view() const noexcept { return const_view(); }
  const_view_type const_view() const noexcept { return {vertices_.const_view(), cells_.const_view()}; }
  index_t num_vertices() const noexcept { return vertices_.shape(0); }
  index_t num_cells() const noexcept { return cells_.shape(0); }
  auto vertices() noexcept { return vertices_.view(); }
  auto cells() noexcept { return cells_.view(); }
  auto vertices() const noexcept { return vertices_.const_view(); }
  auto cells() const noexcept { return cells_.const_view(); }

  // a copy of the mesh on another device
  template <typename Device2>
  basic_unstructured<Scalar, Device2, PhysicalDim, TopologyDim> to(Device2 = {}) const {
    basic_unstructured<Scalar, Device2, PhysicalDim, TopologyDim> out(num_vertices(), num_cells());
    mp::copy(out.vertices(), vertices());
    mp::copy(out.cells(), cells());
    return out;
  }

 private:
  batched_vertex_buffer vertices_;
  batched_cell_buffer cells_;
};

template <typename Scalar, typename Device>
using tetmesh_view = basic_unstructured_view<Scalar, Device, 3, 4>;
template <typename Scalar, typename Device>
using tet_mesh = basic_unstructured<Scalar, Device, 3, 4>;
template <typename Scalar, typename Device>
using tri_mesh = basic_unstructured<Scalar, Device, 2, 3>;

}  // namespace ssim
