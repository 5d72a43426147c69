// time_step_driver.hip -- steps a stretched unit box through ssim::fem::basic_time_step<double, device::hip, ...>
// (include/ssim/...) with the call sequence of the reference's device driver (tests/fem_cu/main.cu:22-102):
// mesh::unit_box + an affine stretch applied with par::seq().vmap; the time step built from mesh.to(device);
// the clamped face marked by a __device__ lambda through parallel().run on dof_type(); gravity by
// add_ext_force_dof; set_threshold; reset(stepper); the mass summed by visiting mass_matrix(); update_hessian();
// then per step prepare_step, update_energy_and_gradients, compute_step(stepper), step_next.
//
// usage: time_step_driver nx steps stepper out.txt [final.npy]
//   mesh unit_box(4·nx, nx, nx) stretched ×4 along x; dt 1e-2, E 1e6, ν 0.4, ρ 10; the first nx·nx vertices
//   (x = 0) clamped; stepper newton | newton_plain (CG without preconditioner) | lbfgs | ncg | forward_euler.
//   out.txt: "mass <Σ M / 3>", "threshold <abs>", then per step "<step> <energy> <residual> <iterations>
//   <converged>"; final.npy: the positions X + u after the last step, [nv, 3] float64.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

#include "mathprim/blas/hip.hpp"
#include "mathprim/core/devices/hip.hpp"
#include "mathprim/core/functional.hpp"
#include "mathprim/linalg/iterative/precond/diagonal.hpp"
#include "mathprim/linalg/iterative/solver/cg.hpp"
#include "mathprim/parallel/hip.hpp"
#include "mathprim/sparse/blas/hip.hpp"
#include "mathprim/supports/io/npy.hpp"
#include "ssim/elast/stable_neohookean.hpp"
#include "ssim/finite_elements/stepper/backward_euler.hpp"
#include "ssim/finite_elements/stepper/nonlinear_cg.hpp"
#include "ssim/finite_elements/time_step.hpp"
#include "ssim/mesh/common_shapes.hpp"

using namespace ssim;
using namespace mathprim;

using Scalar = double;
using ElastModel = elast::stable_neohookean<Scalar, device::hip, 3>;
using SparseBlas = sparse::blas::hip<Scalar, sparse::sparse_format::csr>;
using Blas = blas::hip<Scalar>;
using Step = fem::basic_time_step<Scalar, device::hip, 3, 4, ElastModel, SparseBlas, Blas, par::hip>;
using Jacobi = sparse::iterative::diagonal_preconditioner<Scalar, device::hip, sparse::sparse_format::csr, Blas>;
using JacobiCG = sparse::iterative::cg<Scalar, device::hip, SparseBlas, Blas, Jacobi>;
using PlainCG = sparse::iterative::cg<Scalar, device::hip, SparseBlas, Blas>;

template <typename Solver>
static int run(index_t nx, int steps, Solver& solver, const char* out_path, const char* npy_path) {
  auto mesh = mesh::unit_box<Scalar>(4 * nx, nx, nx);
  functional::affine_transform<Scalar, device::cpu, 3> stretch;
  stretch.lin_[0] = 4;
  par::seq().vmap(stretch, mesh.vertices());

  Step step{mesh.to(device::hip()), 1e-2, 1e6, 0.4, 10};
  step.parallel().run(nx * nx, [dof = step.dof_type()] __device__(index_t i) {
    for (index_t c = 0; c < 3; ++c) dof(i, c) = fem::node_boundary_type::dirichlet;
  });
  step.add_ext_force_dof(1, -9.8);
  step.set_threshold(1e-3);
  step.reset(solver);

  Scalar total = 0;
  auto visit = step.mass_matrix().visit([&total](index_t, index_t, auto& value) { total += value; });
  par::seq().run(visit);
  step.update_hessian();

  std::FILE* out = std::fopen(out_path, "w");
  if (!out) throw std::runtime_error("cannot open output");
  std::fprintf(out, "mass %.17g\nthreshold %.17g\n", total / 3, step.grad_convergence_threshold_abs());
  for (int i = 0; i < steps; ++i) {
    step.prepare_step();
    const Scalar energy = step.update_energy_and_gradients();
    const Scalar residual = step.compute_step(solver);
    step.step_next();
    std::fprintf(out, "%d %.17g %.17g %d %d\n", i, energy, residual, static_cast<int>(solver.result.iterations_),
                 solver.result.converged_ ? 1 : 0);
  }
  std::fclose(out);

  if (npy_path) {  // positions X + u, as the reference drivers' deform_<step>.npy frames
    auto x = make_buffer<Scalar, device::hip>(step.deformation().shape());
    copy(x.view(), step.deformation());
    step.blas().axpy(1.0, step.mesh().vertices(), x.view());
    auto host = make_buffer<Scalar>(x.shape());
    copy(host.view(), x.view());
    std::ofstream f(npy_path, std::ios::binary);
    io::numpy<Scalar, 2>().write(f, host.view());
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc != 5 && argc != 6) {
    std::fprintf(stderr, "usage: %s nx steps newton|newton_plain|lbfgs|ncg|forward_euler out.txt [final.npy]\n",
                 argv[0]);
    return 2;
  }
  try {
    const index_t nx = std::atoi(argv[1]);
    const int steps = std::atoi(argv[2]);
    const std::string method = argv[3];
    const char* npy = argc == 6 ? argv[5] : nullptr;
    if (method == "newton") {
      fem::time_step_solver_backward_euler<JacobiCG> s;
      return run(nx, steps, s, argv[4], npy);
    }
    if (method == "newton_plain") {
      fem::time_step_solver_backward_euler<PlainCG> s;
      return run(nx, steps, s, argv[4], npy);
    }
    if (method == "lbfgs") {
      fem::time_step_solver_lbfgs s;
      return run(nx, steps, s, argv[4], npy);
    }
    if (method == "ncg") {
      fem::time_step_solver_ncg s;
      return run(nx, steps, s, argv[4], npy);
    }
    if (method == "forward_euler") {
      fem::time_step_solver_forward_euler s;
      return run(nx, steps, s, argv[4], npy);
    }
    std::fprintf(stderr, "unknown stepper %s\n", method.c_str());
    return 2;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
