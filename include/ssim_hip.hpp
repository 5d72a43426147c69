// ssim_hip.hpp -- header-only C++17 front end of libssim_hip.so (the C ABI in ssim_hip.h).
//
// ssim::hip::time_step has the member functions of the reference's ssim::fem::basic_time_step
// (finite_elements/time_step.hpp:168-715) with the same argument meaning, so that a binding or driver written
// against basic_time_step<double, device::cuda, ...> ports by changing the type: reset / prepare_step /
// compute_step / step_next, update_energy_and_gradients / update_hessian / check_convergence, add_ext_force_dof,
// restore, the setters (which throw std::invalid_argument where the reference does, time_step.hpp:316-347) and
// the vertex-array views (device pointers of [nv, 3] row-major doubles, time_step.hpp:275-311).  Steppers are the
// tag types newton_solver (time_step_solver_backward_euler, stepper/backward_euler.hpp:9-41), lbfgs_solver
// (time_step_solver_lbfgs, backward_euler.hpp:43-67), lbfgs_pd_solver, lbfgs_h0_solver, ncg_solver,
// ncg_ext_prec_solver and forward_euler_solver, passed to reset / compute_step as in the reference.
//
// ssim::hip::batch is the batched extension (many meshes per set of launches) and has no reference counterpart.
#ifndef SSIM_HIP_HPP
#define SSIM_HIP_HPP

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "ssim_hip.h"

namespace ssim::hip {

// Status → exception: invalid arguments as std::invalid_argument (the reference's setters and reset), anything
// else as std::runtime_error carrying ssim_last_error().
class error : public std::runtime_error {
 public:
  error(ssim_status_t code, const std::string& what) : std::runtime_error(what), code_(code) {}
  ssim_status_t code() const noexcept { return code_; }

 private:
  ssim_status_t code_;
};

inline void check(ssim_status_t st) {
  if (st == SSIM_OK) return;
  const char* msg = ssim_last_error();
  if (st == SSIM_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg ? msg : "invalid argument");
  throw error(st, msg ? msg : "ssim error");
}

// Non-owning view of a device vertex array: data[3*v + c], v < rows.
struct vertex_view {
  double* data = nullptr;
  int rows = 0;
  double* operator[](int v) const noexcept { return data + 3 * static_cast<long long>(v); }
  long long numel() const noexcept { return 3LL * rows; }
};

// Device view of the mesh (rest positions [nv,3] and cells [nt,4]), owned by the time_step.
struct mesh_view {
  const double* vertices = nullptr;
  const int* cells = nullptr;
  int nv = 0, nt = 0;
};

// Host copy of a scalar CSR matrix (rows and columns ascending; the layout of the reference's sys_matrix views).
struct csr_matrix {
  int rows = 0;
  std::vector<int> outer_ptrs, inner_indices;
  std::vector<double> values;
  long long nnz() const noexcept { return static_cast<long long>(values.size()); }
};

struct newton_solver {
  static constexpr ssim_stepper_t kind = SSIM_STEPPER_NEWTON;
};
struct lbfgs_solver {
  static constexpr ssim_stepper_t kind = SSIM_STEPPER_LBFGS;
};
struct ncg_solver {  // time_step_solver_ncg (stepper/nonlinear_cg.hpp:9-37)
  static constexpr ssim_stepper_t kind = SSIM_STEPPER_NCG;
};
struct forward_euler_solver {  // time_step_solver_forward_euler (finite_elements/time_step.hpp:717-736)
  static constexpr ssim_stepper_t kind = SSIM_STEPPER_FORWARD_EULER;
};
// time_step_solver_pd (stepper/projective_dynamics.hpp:9-88), bound as tet_lbfgs_pd
struct lbfgs_pd_solver {
  static constexpr ssim_stepper_t kind = SSIM_STEPPER_LBFGS_PD;
};
// time_step_solver_h0 (stepper/projective_dynamics.hpp:91-148)
struct lbfgs_h0_solver {
  static constexpr ssim_stepper_t kind = SSIM_STEPPER_LBFGS_H0;
};
// time_step_solver_ncg_with_ext_prec (stepper/nonlinear_cg.hpp:41-97); L from time_step::set_matrix
struct ncg_ext_prec_solver {
  static constexpr ssim_stepper_t kind = SSIM_STEPPER_NCG_EXT_PREC;
};

class time_step {
 public:
  // basic_time_step(mesh, dt, E, ν, ρ) (time_step.hpp:260-265); vertices [nv,3], cells [nt,4] on the host
  time_step(const double* vertices, int nv, const int* cells, int nt, double dt = 1e-3, double youngs = 1e6,
            double poisson = 0.33, double density = 1e3) {
    check(ssim_create(&ctx_, vertices, nv, cells, nt, dt, youngs, poisson, density));
  }
  time_step(const time_step&) = delete;
  time_step& operator=(const time_step&) = delete;
  time_step(time_step&& o) noexcept : ctx_(std::exchange(o.ctx_, nullptr)) {}
  time_step& operator=(time_step&& o) noexcept {
    if (this != &o) {
      ssim_destroy(ctx_);
      ctx_ = std::exchange(o.ctx_, nullptr);
    }
    return *this;
  }
  ~time_step() { ssim_destroy(ctx_); }

  ssim_ctx* handle() const noexcept { return ctx_; }
  void set_stream(void* hip_stream) { check(ssim_set_stream(ctx_, hip_stream)); }

  // lifecycle (time_step.hpp:363-543)
  void reset() { check(ssim_reset(ctx_)); }
  template <typename Solver>
  void reset(Solver&) {
    check(ssim_set_stepper(ctx_, Solver::kind));
    reset();
  }
  void prepare_step() { check(ssim_prepare_step(ctx_)); }
  template <typename Solver>
  double compute_step(Solver&) {
    check(ssim_set_stepper(ctx_, Solver::kind));
    double r = 0;
    check(ssim_compute_step(ctx_, &r));
    return r;
  }
  void step_next() { check(ssim_step_next(ctx_)); }
  void restore() { check(ssim_restore(ctx_)); }

  // the operators a stepper uses (time_step.hpp:577-664)
  double update_energy_and_gradients(bool /*is_deform_grad_updated*/ = false) {
    double e = 0;
    check(ssim_update_energy_and_gradients(ctx_, &e));
    return e;
  }
  void update_hessian(bool /*is_deform_grad_updated*/ = false, bool make_spsd = false, bool filter_dbc = true) {
    check(ssim_update_hessian(ctx_, make_spsd ? 1 : 0, filter_dbc ? 1 : 0));
  }
  bool check_convergence() {
    int c = 0;
    check(ssim_check_convergence(ctx_, &c));
    return c != 0;
  }
  void add_ext_force_dof(int dof_idx, double value) { check(ssim_add_ext_force_dof(ctx_, dof_idx, value)); }

  // boundary conditions (bindings/src/fem_host.cpp:123-144): targets [count,3] on the host
  void mark_dirichlet(const int* verts, int count, const double* targets) {
    check(ssim_mark_dirichlet(ctx_, verts, count, targets));
  }
  void mark_general(const int* verts, int count) { check(ssim_mark_general(ctx_, verts, count)); }

  // setters: validated like the reference (time_step.hpp:316-347); return *this for chaining
  time_step& set_time_step(double v) { return chk(ssim_set_time_step(ctx_, v)); }
  time_step& set_youngs(double v) { return chk(ssim_set_youngs(ctx_, v)); }
  time_step& set_poisson(double v) { return chk(ssim_set_poisson(ctx_, v)); }
  time_step& set_density(double v) { return chk(ssim_set_density(ctx_, v)); }
  time_step& set_damping(double v) { return chk(ssim_set_damping(ctx_, v)); }
  time_step& set_threshold(double v) { return chk(ssim_set_threshold(ctx_, v)); }
  time_step& set_solver_params(const ssim_solver_params& p) { return chk(ssim_set_solver_params(ctx_, &p)); }
  // tet_ncg_cuda_ext_ai::set_matrix (bindings/src/fem_cuda.cu:204): L in the CSR order of the system matrix
  time_step& set_matrix(const double* values, long long nnz) { return chk(ssim_set_precond_matrix(ctx_, values, nnz)); }
  // the ElastModel template argument of basic_time_step, chosen at run time
  time_step& set_material(ssim_material_t model) { return chk(ssim_set_material_model(ctx_, model)); }
  // getters (time_step.hpp:312-313, 349-353, 361); the reference's time_step() is time_step_size() here, since
  // a member cannot take the class's name
  double time_step_size() const { return param(0); }
  double youngs() const { return param(1); }
  double poisson() const { return param(2); }
  double density() const { return param(3); }
  double damping() const { return param(4); }
  double grad_convergence_threshold() const { return param(5); }
  long long total_dofs() const { return 3LL * num_vertices(); }
  // update_deformation_gradient() (time_step.hpp:571-575): F is formed inside every element kernel from the
  // current next_deform, so there is nothing to precompute; kept so that reference call sites compile
  void update_deformation_gradient() const noexcept {}
  ssim_solver_params solver_params() const {
    ssim_solver_params p{};
    check(ssim_get_solver_params(ctx_, &p));
    return p;
  }

  // vertex-array views on the device (time_step.hpp:275-311)
  vertex_view deformation() const { return view(SSIM_ARRAY_DEFORMATION); }
  vertex_view prev_deform() const { return view(SSIM_ARRAY_PREV_DEFORM); }
  vertex_view inertia_deform() const { return view(SSIM_ARRAY_INERTIA); }
  vertex_view next_deform() const { return view(SSIM_ARRAY_NEXT_DEFORM); }
  vertex_view velocity() const { return view(SSIM_ARRAY_VELOCITY); }
  vertex_view ext_accel() const { return view(SSIM_ARRAY_EXT_ACCEL); }
  vertex_view forces() const { return view(SSIM_ARRAY_FORCES); }
  vertex_view dbc_values() const { return view(SSIM_ARRAY_DBC_VALUES); }

  // host copies
  std::vector<double> get(ssim_array_t which) const {
    std::vector<double> out(3 * static_cast<size_t>(num_vertices()));
    check(ssim_get_array(ctx_, which, out.data()));
    return out;
  }
  void set(ssim_array_t which, const std::vector<double>& values) {
    if (values.size() != 3 * static_cast<size_t>(num_vertices())) throw std::invalid_argument("array size mismatch");
    check(ssim_set_array(ctx_, which, values.data()));
  }

  int num_vertices() const {
    int n = 0;
    check(ssim_num_vertices(ctx_, &n));
    return n;
  }
  int num_cells() const {
    int n = 0;
    check(ssim_num_cells(ctx_, &n));
    return n;
  }
  double total_mass() const {
    double m = 0;
    check(ssim_total_mass(ctx_, &m));
    return m;
  }
  double grad_convergence_threshold_abs() const {
    double t = 0;
    check(ssim_grad_threshold_abs(ctx_, &t));
    return t;
  }
  ssim_step_stats last_stats() const {
    ssim_step_stats s{};
    check(ssim_get_step_stats(ctx_, &s));
    return s;
  }

  // the remaining views of time_step.hpp:275-311, as device views (mesh) or host copies (the rest)
  mesh_view mesh() const {
    mesh_view m;
    check(ssim_device_mesh(ctx_, &m.vertices, &m.cells));
    m.nv = num_vertices();
    m.nt = num_cells();
    return m;
  }
  std::vector<unsigned char> dof_type() const {
    std::vector<unsigned char> out(3 * static_cast<size_t>(num_vertices()));
    check(ssim_get_dof_type(ctx_, out.data()));
    return out;
  }
  std::vector<double> dminv() const {  // [nt,3,3] row-major
    std::vector<double> out(9 * static_cast<size_t>(num_cells()));
    check(ssim_get_rest_shape(ctx_, out.data(), nullptr));
    return out;
  }
  std::vector<double> rest_vol() const {  // [nt]
    std::vector<double> out(static_cast<size_t>(num_cells()));
    check(ssim_get_rest_shape(ctx_, nullptr, out.data()));
    return out;
  }
  // local_stiff() [nt,12,12]: pfpxᵀ·H·pfpx at next_deform (time_step.hpp:641-644), computed on request -- the
  // Hessian kernels assemble it without storing it
  std::vector<double> local_stiff(bool make_spsd = true) const {
    const std::vector<double> u = get(SSIM_ARRAY_NEXT_DEFORM);
    std::vector<double> out(144 * static_cast<size_t>(num_cells()));
    check(ssim_eval_elements(ctx_, u.data(), make_spsd ? 1 : 0, nullptr, nullptr, nullptr, out.data()));
    return out;
  }
  csr_matrix sysmatrix() const { return csr(ssim_export_hessian_csr); }
  csr_matrix mass_matrix() const { return csr(ssim_export_mass_csr); }
  csr_matrix mass_matrix_filtered() const { return csr(ssim_export_mass_filtered_csr); }

 private:
  csr_matrix csr(ssim_status_t (*exporter)(ssim_ctx*, int*, int*, double*)) const {
    csr_matrix m;
    long long nnz = 0;
    check(ssim_csr_nnz(ctx_, &nnz));
    m.rows = 3 * num_vertices();
    m.outer_ptrs.resize(static_cast<size_t>(m.rows) + 1);
    m.inner_indices.resize(static_cast<size_t>(nnz));
    m.values.resize(static_cast<size_t>(nnz));
    check(exporter(ctx_, m.outer_ptrs.data(), m.inner_indices.data(), m.values.data()));
    return m;
  }
  double param(int which) const {
    double v[6] = {};
    check(ssim_get_physical_params(ctx_, &v[0], &v[1], &v[2], &v[3], &v[4], &v[5]));
    return v[which];
  }
  time_step& chk(ssim_status_t st) {
    check(st);
    return *this;
  }
  vertex_view view(ssim_array_t which) const {
    vertex_view v;
    check(ssim_device_array(ctx_, which, &v.data));
    v.rows = num_vertices();
    return v;
  }
  ssim_ctx* ctx_ = nullptr;
};

// Batched extension: meshes given as per-mesh counts and concatenated arrays (see ssim_batch_create).
class batch {
 public:
  batch(int nmesh, const int* nv, const int* nt, const double* vertices, const int* cells, const double* lambda,
        const double* mu, double density, double dt) {
    check(ssim_batch_create(&b_, nmesh, nv, nt, vertices, cells, lambda, mu, density, dt));
  }
  batch(const batch&) = delete;
  batch& operator=(const batch&) = delete;
  batch(batch&& o) noexcept : b_(std::exchange(o.b_, nullptr)) {}
  ~batch() { ssim_batch_destroy(b_); }

  ssim_batch* handle() const noexcept { return b_; }
  void reset() { check(ssim_batch_reset(b_)); }
  void step() { check(ssim_batch_step(b_)); }
  void set_threshold(double rtol) { check(ssim_batch_set_threshold(b_, rtol)); }
  void mark_dirichlet(int mesh, const int* verts, int count, const double* targets) {
    check(ssim_batch_mark_dirichlet(b_, mesh, verts, count, targets));
  }
  void set(ssim_array_t which, int mesh, const double* host) { check(ssim_batch_set_array(b_, which, mesh, host)); }
  void get(ssim_array_t which, int mesh, double* host) const { check(ssim_batch_get_array(b_, which, mesh, host)); }
  ssim_step_stats last_stats(int mesh = -1) const {
    ssim_step_stats s{};
    check(ssim_batch_get_step_stats(b_, mesh, &s));
    return s;
  }

 private:
  ssim_batch* b_ = nullptr;
};

}  // namespace ssim::hip

#endif
