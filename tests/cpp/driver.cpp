// driver.cpp -- a C++ program stepping a bar through the header-only front end (include/ssim_hip.hpp) the way
// the reference's drivers use basic_time_step (tests/fem_time_step/main.cpp): construct from a mesh, mark the
// clamped vertices, add gravity, reset with a stepper, then prepare_step / compute_step / step_next.
//
// Input file (little-endian): int32 nv, nt, steps, stepper (0 Newton, 1 L-BFGS, 2 NCG, 3 forward Euler); float64 dt, E, nu, rho, rtol,
// gravity_y; float64 X[nv*3]; int32 cells[nt*4]; uint8 clamped[nv]; float64 u0[nv*3].
// Output file: float64 u[nv*3] after the last step, then per step int32 newton_iterations, pcg_iterations, then
// float64 sums of absolute values over the remaining views: rest_vol, dminv, sysmatrix nnz and values, mass_matrix,
// mass_matrix_filtered, dof_type, local_stiff.
// With a third argument, a folder, the driver also writes the positions X + u after every step to
// <folder>/deform_<step>.npy as an [nv, 3] float64 array, the frames tests/fem_cu/main.cu:86-92 writes and
// tests/fem_time_step/visualize.py:9-12 reads.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <vector>

#include "ssim_hip.hpp"
#include "ssim_npy.hpp"

template <typename T>
static void rd(FILE* f, T* p, size_t n) {
  if (fread(p, sizeof(T), n, f) != n) throw std::runtime_error("short input file");
}

int main(int argc, char** argv) {
  if (argc != 3 && argc != 4) {
    std::fprintf(stderr, "usage: %s input output [frames_folder]\n", argv[0]);
    return 2;
  }
  try {
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) throw std::runtime_error("cannot open input");
    int32_t hdr[4];
    double prm[6];
    rd(in, hdr, 4);
    rd(in, prm, 6);
    const int nv = hdr[0], nt = hdr[1], steps = hdr[2], stepper = hdr[3];
    std::vector<double> X(3 * (size_t)nv), u0(3 * (size_t)nv);
    std::vector<int32_t> cells(4 * (size_t)nt);
    std::vector<uint8_t> clamped(nv);
    rd(in, X.data(), X.size());
    rd(in, cells.data(), cells.size());
    rd(in, clamped.data(), clamped.size());
    rd(in, u0.data(), u0.size());
    std::fclose(in);

    ssim::hip::time_step step(X.data(), nv, cells.data(), nt, prm[0], prm[1], prm[2], prm[3]);
    std::vector<int> fixed;
    for (int v = 0; v < nv; ++v)
      if (clamped[v]) fixed.push_back(v);
    std::vector<double> zeros(3 * fixed.size(), 0.0);
    step.mark_dirichlet(fixed.data(), (int)fixed.size(), zeros.data());
    step.add_ext_force_dof(1, prm[5]);
    step.set(SSIM_ARRAY_DEFORMATION, u0);
    step.set_threshold(prm[4]);
    ssim::hip::newton_solver newton;
    ssim::hip::lbfgs_solver lbfgs;
    ssim::hip::ncg_solver ncg;
    ssim::hip::forward_euler_solver euler;
    if (stepper == 1)
      step.reset(lbfgs);
    else if (stepper == 2)
      step.reset(ncg);
    else if (stepper == 3)
      step.reset(euler);
    else
      step.reset(newton);
    std::vector<int32_t> counts;
    for (int s = 0; s < steps; ++s) {
      step.prepare_step();
      if (stepper == 1)
        step.compute_step(lbfgs);
      else if (stepper == 2)
        step.compute_step(ncg);
      else if (stepper == 3)
        step.compute_step(euler);
      else
        step.compute_step(newton);
      step.step_next();
      const ssim_step_stats st = step.last_stats();
      counts.push_back(st.newton_iterations);
      counts.push_back(st.pcg_iterations);
      if (argc == 4) {
        std::vector<double> x = step.get(SSIM_ARRAY_DEFORMATION);
        for (size_t i = 0; i < x.size(); ++i) x[i] += X[i];
        ssim::hip::io::write_npy(std::string(argv[3]) + "/deform_" + std::to_string(s) + ".npy", x.data(),
                                 (size_t)nv, 3);
      }
    }
    const std::vector<double> u = step.get(SSIM_ARRAY_DEFORMATION);
    // the setters reject what the reference rejects
    bool threw = false;
    try {
      step.set_poisson(0.5);
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    if (!threw) throw std::runtime_error("set_poisson(0.5) did not throw");
    // the getters return what was set (time_step.hpp:349-353)
    if (step.time_step_size() != prm[0] || step.youngs() != prm[1] || step.poisson() != prm[2] ||
        step.density() != prm[3] || step.grad_convergence_threshold() != prm[4] || step.damping() != 0.0 ||
        step.total_dofs() != 3LL * nv)
      throw std::runtime_error("getter mismatch");
    step.update_deformation_gradient();
    // the remaining views of basic_time_step (time_step.hpp:275-311)
    const ssim::hip::mesh_view mv = step.mesh();
    if (mv.nv != nv || mv.nt != nt || !mv.vertices || !mv.cells) throw std::runtime_error("mesh() view mismatch");
    auto sum = [](const auto& v) {
      double t = 0;
      for (auto x : v) t += std::fabs(static_cast<double>(x));
      return t;
    };
    const ssim::hip::csr_matrix H = step.sysmatrix();
    const double views[8] = {sum(step.rest_vol()), sum(step.dminv()), static_cast<double>(H.nnz()), sum(H.values),
                             sum(step.mass_matrix().values), sum(step.mass_matrix_filtered().values),
                             sum(step.dof_type()), sum(step.local_stiff())};
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) throw std::runtime_error("cannot open output");
    std::fwrite(u.data(), sizeof(double), u.size(), out);
    std::fwrite(counts.data(), sizeof(int32_t), counts.size(), out);
    std::fwrite(views, sizeof(double), 8, out);
    std::fclose(out);
    std::printf("driver: %d vertices, %d tets, %d steps done\n", nv, nt, steps);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "driver failed: %s\n", e.what());
    return 1;
  }
  return 0;
}
