// ssim/mesh/common_shapes.hpp -- host mesh generators of the reference drivers (ssim/include/ssim/mesh/
// common_shapes.hpp:59-145): unit_box(nx, ny, nz) over [0, 1]³ with nx·ny·nz vertices (x-slowest ids) and five
// tets per grid cube, the cube's corner labels mirrored along each axis whose cube index is odd so that
// neighbouring cubes share their face diagonals; tetra(), the single reference tetrahedron.
#pragma once

#include "ssim/mesh/basic_mesh.hpp"

namespace ssim::mesh {

template <typename Scalar>
tet_mesh<Scalar, mp::device::cpu> unit_box(index_t nx = 2, index_t ny = 2, index_t nz = 2) {
  if (nx < 2 || ny < 2 || nz < 2) throw std::invalid_argument("unit_box needs at least 2 vertices per axis");
  tet_mesh<Scalar, mp::device::cpu> mesh(nx * ny * nz, 5 * (nx - 1) * (ny - 1) * (nz - 1));
  auto X = mesh.vertices();
  auto C = mesh.cells();
  const auto vid = [&](index_t i, index_t j, index_t k) { return (i * ny + j) * nz + k; };
  for (index_t i = 0; i < nx; ++i)
    for (index_t j = 0; j < ny; ++j)
      for (index_t k = 0; k < nz; ++k) {
        const index_t v = vid(i, j, k);
        X(v, 0) = i / Scalar(nx - 1);
        X(v, 1) = j / Scalar(ny - 1);
        X(v, 2) = k / Scalar(nz - 1);
      }
  // the five tets of a cube by corner label (bit 2: x, bit 1: y, bit 0: z), first vertex first
  static constexpr int kTets[5][4] = {{0b000, 0b100, 0b010, 0b001},
                                      {0b010, 0b100, 0b110, 0b111},
                                      {0b100, 0b010, 0b001, 0b111},
                                      {0b100, 0b001, 0b101, 0b111},
                                      {0b011, 0b001, 0b010, 0b111}};
  index_t t = 0;
  for (index_t i = 0; i + 1 < nx; ++i)
    for (index_t j = 0; j + 1 < ny; ++j)
      for (index_t k = 0; k + 1 < nz; ++k) {
        const int flip = ((i & 1) << 2) | ((j & 1) << 1) | (k & 1);  // mirrored axes of this cube
        const auto corner = [&](int label) {
          const int c = label ^ flip;
          return vid(i + ((c >> 2) & 1), j + ((c >> 1) & 1), k + (c & 1));
        };
        for (const auto& tet : kTets) {
          for (int a = 0; a < 4; ++a) C(t, a) = corner(tet[a]);
          ++t;
        }
      }
  return mesh;
}

template <typename Scalar>
tet_mesh<Scalar, mp::device::cpu> tetra() {
  tet_mesh<Scalar, mp::device::cpu> mesh(4, 1);
  auto X = mesh.vertices();
  for (index_t v = 0; v < 4; ++v)
    for (index_t c = 0; c < 3; ++c) X(v, c) = (v == c + 1) ? Scalar(1) : Scalar(0);
  for (index_t a = 0; a < 4; ++a) mesh.cells()(0, a) = a;
  return mesh;
}

}  // namespace ssim::mesh
