// host_api_check.cpp -- the host side of the header-only API (include/mathprim, include/ssim/mesh), compiled by
// g++ without HIP: unit_box / tetra, views (batched rows, element access, flatten, reshape), make_buffer, copy,
// par::seq run / vmap with affine_transform, a CSR visit task and the .npy writer.
//
// usage: host_api_check nx ny nz out.bin out.npy
//   out.bin: int32 nv, nt; float64 X[nv·3] after stretching x by 4; int32 cells[nt·4]; float64 checks[4]
//   (Σ visited CSR values, Σ of the flattened view, reshape(3, nv)(1, 0), row-view sum of vertex 1)
//   out.npy: X as an [nv, 3] float64 array
#include <cstdio>
#include <cstdlib>
#include <fstream>

#include "mathprim/core/buffer.hpp"
#include "mathprim/parallel/parallel.hpp"
#include "mathprim/sparse/basic_sparse.hpp"
#include "mathprim/supports/io/npy.hpp"
#include "ssim/mesh/common_shapes.hpp"

using namespace mathprim;

int main(int argc, char** argv) {
  if (argc != 6) return 2;
  auto mesh = ssim::mesh::unit_box<double>(std::atoi(argv[1]), std::atoi(argv[2]), std::atoi(argv[3]));
  functional::affine_transform<double, device::cpu, 3> stretch;
  stretch.lin_[0] = 4;
  par::seq().vmap(stretch, mesh.vertices());

  auto copy_buf = make_buffer<double>(mesh.vertices().shape());
  copy(copy_buf.view(), mesh.vertices());
  double flat = 0;
  auto f = copy_buf.view().flatten();
  par::seq().run(f.numel(), [&](index_t i) { flat += f[i]; });
  const double reshaped = copy_buf.view().reshape(3, mesh.num_vertices())(1, 0);
  auto row = mesh.vertices()[1];
  const double rowsum = row[0] + row[1] + row[2];

  // a 2×2 CSR [[1, 2], [0, 3]] visited on the host
  sparse::basic_sparse_matrix<double, device::cpu, sparse::sparse_format::csr> m(2, 2, 3);
  const int outer[3] = {0, 2, 3}, inner[3] = {0, 1, 1};
  const double vals[3] = {1, 2, 3};
  for (int i = 0; i < 3; ++i) {
    m.outer_data()[i] = outer[i];
    m.inner_data()[i] = inner[i];
    m.values_data()[i] = vals[i];
  }
  double visited = 0;
  auto task = m.view().visit([&](index_t r, index_t c, auto& v) { visited += v * (1 + r + 10 * c); });
  par::seq().run(task);

  std::FILE* out = std::fopen(argv[4], "wb");
  if (!out) return 1;
  const int hdr[2] = {mesh.num_vertices(), mesh.num_cells()};
  std::fwrite(hdr, sizeof(int), 2, out);
  std::fwrite(mesh.vertices().data(), sizeof(double), 3 * static_cast<size_t>(hdr[0]), out);
  std::fwrite(mesh.cells().data(), sizeof(int), 4 * static_cast<size_t>(hdr[1]), out);
  const double checks[4] = {visited, flat, reshaped, rowsum};
  std::fwrite(checks, sizeof(double), 4, out);
  std::fclose(out);
  std::ofstream npy(argv[5], std::ios::binary);
  io::numpy<double, 2>().write(npy, mesh.vertices());
  return 0;
}
