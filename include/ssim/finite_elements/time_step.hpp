// ssim/finite_elements/time_step.hpp -- ssim::fem::basic_time_step<Scalar, Device, PhysicalDim, TopologyDim,
// ElastModel, SparseBlas, Blas, ParImpl> for device::hip: the reference's header-only time step
// (finite_elements/time_step.hpp:103-715) with the same members, argument meaning and exceptions, whose work runs
// in libssim_hip.so through the C ABI of ssim_hip.h.
//
// Instantiation: Scalar = double, Device = mp::device::hip, PhysicalDim = 3, TopologyDim = 4, ElastModel =
// elast::stable_neohookean<double, device::hip, 3> or elast::linear<...>, SparseBlas = mp::sparse::blas::hip<...>,
// Blas = mp::blas::hip<double>, ParImpl = mp::par::hip (the HIP counterparts of the reference's cusparse, cublas
// and par::cuda arguments, tests/fem_cu/main.cu:31-34).
//
// What lives where:
// * state vectors (deformation, velocity, forces, ...) and the Dirichlet flags are device views of the
//   library's own arrays ([nv, 3] row-major doubles / bytes), so kernels and __device__ lambdas read and write
//   them in place (tests/fem_cu/main.cu:41-45);
// * the rest data and matrices the reference keeps on the device in its own layouts (dminv, rest_vol,
//   local_stiff, sysmatrix, mass_matrix, mass_matrix_filtered) are returned as host copies in the reference's
//   layouts, refreshed on each call: the library stores Dm⁻¹ component-major and the system matrix as a
//   symmetric blocked ELL, and never materialises the per-tet 12×12 stiffness;
// * ElastModel selects the element kernels; a per-call material object is not used (the reference's
//   update_* build their model from (E, ν) too, time_step.hpp:595-598, 633-635).
#pragma once

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "mathprim/blas/hip.hpp"
#include "mathprim/optim/basic_optim.hpp"
#include "mathprim/parallel/hip.hpp"
#include "mathprim/sparse/blas/hip.hpp"
#include "ssim/elast/basic_elast.hpp"
#include "ssim/finite_elements/boundary.hpp"
#include "ssim/finite_elements/defines.hpp"
#include "ssim/mesh/basic_mesh.hpp"
#include "ssim_hip.h"

namespace ssim::fem {

namespace internal {
// ssim_status_t → exception: std::invalid_argument where the reference throws it (setters, reset), otherwise
// std::runtime_error with ssim_last_error()
inline void check(ssim_status_t st) {
  if (st == SSIM_OK) return;
  const char* msg = ssim_last_error();
  if (st == SSIM_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg ? msg : "invalid argument");
  throw std::runtime_error(msg ? msg : "ssim error");
}
}  // namespace internal

template <typename Scalar, typename Device,                               //
          index_t PhysicalDim, index_t TopologyDim, typename ElastModel,  //
          typename SparseBlas, typename Blas, typename ParImpl>
class basic_time_step;

// CRTP base of the steppers (time_step.hpp:108-129): solve() / reset() forward to solve_impl / reset_impl.
template <typename Derived>
class basic_time_step_solver {
 public:
  using derived_type = Derived;
  Derived& derived() noexcept { return static_cast<Derived&>(*this); }
  const Derived& derived() const noexcept { return static_cast<const Derived&>(*this); }
  template <typename Scalar, typename Device, index_t P, index_t T, typename E, typename SB, typename B, typename Par>
  void solve(basic_time_step<Scalar, Device, P, T, E, SB, B, Par>& step) {
    derived().solve_impl(step);
  }
  template <typename Scalar, typename Device, index_t P, index_t T, typename E, typename SB, typename B, typename Par>
  void reset(basic_time_step<Scalar, Device, P, T, E, SB, B, Par>& step) {
    derived().reset_impl(step);
  }
};

// A stepper of the library: one of the ssim_stepper_t solvers, with optional solver parameters (unset: the
// library's defaults, ≤ 10 Newton iterations, Jacobi-PCG to 1e-8, Armijo 0.3 / 0.5 / 20 trials, L-BFGS m = 8)
// and the record of its last solve (mathprim optim_result, stepper/backward_euler.hpp:40).
template <typename Derived, ssim_stepper_t Kind>
class library_stepper : public basic_time_step_solver<Derived> {
 public:
  static constexpr ssim_stepper_t kind = Kind;
  template <typename Step>
  void solve_impl(Step& s) {
    s.run_library_solve(kind, params_set_ ? &params_ : nullptr, static_cast<Derived&>(*this).precond(), result);
  }
  template <typename Step>
  void reset_impl(Step&) {}
  int precond() const noexcept { return -1; }  // keep the context's PCG preconditioner
  void set_params(const ssim_solver_params& p) noexcept {
    params_ = p;
    params_set_ = true;
  }
  mp::optim::optim_result<double> result;

 private:
  ssim_solver_params params_{};
  bool params_set_ = false;
};

// time_step_solver_nothing (time_step.hpp:131-143): reset only, no solve
class time_step_solver_nothing : public basic_time_step_solver<time_step_solver_nothing> {
 public:
  template <typename Step>
  void solve_impl(Step&) {}
  template <typename Step>
  void reset_impl(Step&) {}
};

// time_step_solver_forward_euler (time_step.hpp:717-736): next_deform = û + ½·dt²·forces(û)
class time_step_solver_forward_euler
    : public library_stepper<time_step_solver_forward_euler, SSIM_STEPPER_FORWARD_EULER> {};

template <typename Scalar, typename Device,                               //
          index_t PhysicalDim, index_t TopologyDim, typename ElastModel,  //
          typename SparseBlas, typename Blas, typename ParImpl>
class basic_time_step {
  static_assert(std::is_same_v<Scalar, double>, "the HIP time step computes in FP64");
  static_assert(std::is_same_v<Device, mp::device::hip>, "the HIP build instantiates basic_time_step on device::hip");
  static_assert(PhysicalDim == 3 && TopologyDim == 4, "tetrahedra in 3D");

 public:
  using scalar_type = Scalar;
  using device_type = Device;
  static constexpr index_t physical_dim = PhysicalDim;
  static constexpr index_t topology_dim = TopologyDim;
  static constexpr index_t dofs_per_node = PhysicalDim;
  static constexpr index_t hessian_nrows = topology_dim * dofs_per_node;
  using blas_ref = mp::blas::basic_blas<Blas, Scalar, Device>&;
  using par_ref = mp::par::parfor<ParImpl>&;

  using mesh_type = basic_unstructured<Scalar, Device, PhysicalDim, TopologyDim>;
  using const_mesh_view = basic_unstructured_view<const Scalar, Device, PhysicalDim, TopologyDim>;
  using mesh_view = const_mesh_view;  // the rest mesh is owned by the library and read-only
  using vertex_type = mp::contiguous_view<Scalar, mp::shape_t<PhysicalDim>, Device>;
  using batched_vertex = mp::batched<vertex_type>;
  using const_batched_vertex = mp::batched<mp::contiguous_view<const Scalar, mp::shape_t<PhysicalDim>, Device>>;
  using dof_type_type = mp::contiguous_view<node_boundary_type, mp::shape_t<PhysicalDim>, Device>;
  using batched_dof_type = mp::batched<dof_type_type>;
  using const_batched_dof_type =
      mp::batched<mp::contiguous_view<const node_boundary_type, mp::shape_t<PhysicalDim>, Device>>;
  // host copies (see the file comment)
  using host_sys_matrix = sparse_matrix<Scalar, mp::device::cpu>;
  using sys_matrix_view = sparse_view<Scalar, mp::device::cpu>;
  using const_sys_matrix_view = sparse_view<const Scalar, mp::device::cpu>;
  using dminv_type = mp::batched<mp::contiguous_view<Scalar, mp::shape_t<PhysicalDim, PhysicalDim>, mp::device::cpu>>;
  using rest_volume = mp::contiguous_vector_view<Scalar, mp::device::cpu>;
  using batched_local_stiffness =
      mp::batched<mp::contiguous_view<Scalar, mp::shape_t<hessian_nrows, hessian_nrows>, mp::device::cpu>>;

  // basic_time_step(mesh, dt, E, ν, ρ) + reset() (time_step.hpp:260-265): the mesh is copied into the library
  explicit basic_time_step(mesh_type mesh, Scalar time_step = 1e-3, Scalar youngs = 1e6, Scalar poisson = 0.33,
                           Scalar density = 1e3) {
    auto host = mesh.to(mp::device::cpu{});
    internal::check(ssim_create(&ctx_, host.vertices().data(), host.num_vertices(), host.cells().data(),
                                host.num_cells(), time_step, youngs, poisson, density));
    internal::check(ssim_set_material_model(ctx_, ElastModel::material));
    nv_ = host.num_vertices();
    nt_ = host.num_cells();
  }
  basic_time_step(const basic_time_step&) = delete;
  basic_time_step& operator=(const basic_time_step&) = delete;
  basic_time_step(basic_time_step&& o) noexcept { *this = std::move(o); }
  basic_time_step& operator=(basic_time_step&& o) noexcept {
    if (this != &o) {
      if (ctx_) ssim_destroy(ctx_);
      ctx_ = o.ctx_;
      o.ctx_ = nullptr;
      nv_ = o.nv_;
      nt_ = o.nt_;
      residual_ = o.residual_;
    }
    return *this;
  }
  ~basic_time_step() {
    if (ctx_) ssim_destroy(ctx_);
  }

  blas_ref blas() noexcept { return blas_; }
  par_ref parallel() noexcept { return pf_; }

  ////////// views (time_step.hpp:275-311) //////////
  const_mesh_view mesh() const {
    const double* X = nullptr;
    const int* C = nullptr;
    internal::check(ssim_device_mesh(ctx_, &X, &C));
    return {{X, mp::make_shape(nv_, mp::holder<3>{})}, {C, mp::make_shape(nt_, mp::holder<4>{})}};
  }
  batched_dof_type dof_type() {
    unsigned char* p = nullptr;
    internal::check(ssim_device_dof_type(ctx_, &p));
    return {reinterpret_cast<node_boundary_type*>(p), mp::make_shape(nv_, mp::holder<3>{})};
  }
  batched_vertex dbc_values() { return array(SSIM_ARRAY_DBC_VALUES); }
  batched_vertex deformation() { return array(SSIM_ARRAY_DEFORMATION); }
  batched_vertex prev_deform() { return array(SSIM_ARRAY_PREV_DEFORM); }
  batched_vertex inertia_deform() { return array(SSIM_ARRAY_INERTIA); }
  batched_vertex next_deform() { return array(SSIM_ARRAY_NEXT_DEFORM); }
  batched_vertex velocity() { return array(SSIM_ARRAY_VELOCITY); }
  batched_vertex ext_accel() { return array(SSIM_ARRAY_EXT_ACCEL); }
  batched_vertex forces() { return array(SSIM_ARRAY_FORCES); }

  // dminv() [nt, 3, 3] and rest_vol() [nt] (time_step.hpp:296-301), host copies
  dminv_type dminv() {
    rest_shape();
    return {h_dminv_.data(), mp::make_shape(nt_, mp::holder<3>{}, mp::holder<3>{})};
  }
  rest_volume rest_vol() {
    rest_shape();
    return {h_vol_.data(), mp::shape_t<mp::keep_dim>(nt_)};
  }
  // local_stiff() [nt, 12, 12] (time_step.hpp:298-299): the per-tet stiffness at next_deform with the make_spsd
  // of the last update_hessian, recomputed on the host's request
  batched_local_stiffness local_stiff() {
    std::vector<double> u(3 * static_cast<size_t>(nv_));
    internal::check(ssim_get_array(ctx_, SSIM_ARRAY_NEXT_DEFORM, u.data()));
    h_stiff_.resize(static_cast<size_t>(nt_) * hessian_nrows * hessian_nrows);
    internal::check(ssim_eval_elements(ctx_, u.data(), last_spsd_ ? 1 : 0, nullptr, nullptr, nullptr, h_stiff_.data()));
    return {h_stiff_.data(), mp::make_shape(nt_, mp::holder<hessian_nrows>{}, mp::holder<hessian_nrows>{})};
  }
  // sysmatrix(), mass_matrix(), mass_matrix_filtered() (time_step.hpp:302-311): scalar CSR, rows and columns
  // ascending, host copies of the current values
  sys_matrix_view sysmatrix() { return export_csr(h_sys_, ssim_export_hessian_csr); }
  sys_matrix_view mass_matrix() { return export_csr(h_mass_, ssim_export_mass_csr); }
  sys_matrix_view mass_matrix_filtered() { return export_csr(h_massf_, ssim_export_mass_filtered_csr); }

  Scalar grad_convergence_threshold() const { return param(5); }
  Scalar grad_convergence_threshold_abs() const {
    double t = 0;
    internal::check(ssim_grad_threshold_abs(ctx_, &t));
    return t;
  }

  ////////// setters (time_step.hpp:316-347): std::invalid_argument on the reference's conditions //////////
  basic_time_step& set_time_step(Scalar v) { return chk(ssim_set_time_step(ctx_, v)); }
  basic_time_step& set_youngs(Scalar v) { return chk(ssim_set_youngs(ctx_, v)); }
  basic_time_step& set_poisson(Scalar v) { return chk(ssim_set_poisson(ctx_, v)); }
  basic_time_step& set_density(Scalar v) { return chk(ssim_set_density(ctx_, v)); }
  basic_time_step& set_damping(Scalar v) { return chk(ssim_set_damping(ctx_, v)); }
  basic_time_step& set_threshold(Scalar v) {
    if (!(0 < v && v < 1)) throw std::invalid_argument("Threshold must be positive.");
    return chk(ssim_set_threshold(ctx_, v));
  }
  Scalar time_step() const { return param(0); }
  Scalar youngs() const { return param(1); }
  Scalar poisson() const { return param(2); }
  Scalar density() const { return param(3); }
  Scalar damping() const { return param(4); }

  // uniform_elasticity() (time_step.hpp:356): the reference builds the model from (E, ν) in its (λ, μ) slots;
  // kept as written, since the projective-dynamics stepper's matrix depends on it
  ElastModel uniform_elasticity() const { return ElastModel{youngs(), poisson()}; }
  index_t total_dofs() const noexcept { return nv_ * dofs_per_node; }

  ////////// precompute and loop stages (time_step.hpp:363-569) //////////
  void reset() {
    time_step_solver_nothing alg;
    reset(alg);
  }
  template <typename Algorithm>
  void reset(basic_time_step_solver<Algorithm>& solver) {
    if constexpr (has_kind<Algorithm>::value) internal::check(ssim_set_stepper(ctx_, Algorithm::kind));
    internal::check(ssim_reset(ctx_));
    solver.reset(*this);
  }
  void prepare_step() { internal::check(ssim_prepare_step(ctx_)); }
  template <typename Algorithm>
  Scalar compute_step(basic_time_step_solver<Algorithm>& stepper) {
    residual_ = std::nan("");
    stepper.solve(*this);
    if (std::isnan(residual_)) {  // a stepper that solved nothing: compute_step still evaluates the forces
      update_energy_and_gradients(false);
      std::vector<double> f(3 * static_cast<size_t>(nv_));
      internal::check(ssim_get_array(ctx_, SSIM_ARRAY_FORCES, f.data()));
      double s = 0;
      for (double x : f) s += x * x;
      residual_ = std::sqrt(s);
    }
    return residual_;
  }
  void step_next() { internal::check(ssim_step_next(ctx_)); }
  void restore() { internal::check(ssim_restore(ctx_)); }
  void add_ext_force_dof(index_t dof_idx, Scalar del) { internal::check(ssim_add_ext_force_dof(ctx_, dof_idx, del)); }

  ////////// the operators a stepper uses (time_step.hpp:571-664) //////////
  // F is formed inside every element kernel from next_deform; there is nothing to precompute
  void update_deformation_gradient() noexcept {}
  Scalar update_energy_and_gradients(bool /*is_deform_grad_updated*/ = false) {
    double e = 0;
    internal::check(ssim_update_energy_and_gradients(ctx_, &e));
    return e;
  }
  void update_hessian(bool /*is_deform_grad_updated*/ = false, bool make_spsd = false, bool filter_dbc = true) {
    internal::check(ssim_update_hessian(ctx_, make_spsd ? 1 : 0, filter_dbc ? 1 : 0));
    last_spsd_ = make_spsd;
  }
  bool check_convergence() {
    int c = 0;
    internal::check(ssim_check_convergence(ctx_, &c));
    return c != 0;
  }

  ////////// beyond the reference //////////
  ssim_ctx* handle() const noexcept { return ctx_; }
  ssim_step_stats last_stats() const {
    ssim_step_stats s{};
    internal::check(ssim_get_step_stats(ctx_, &s));
    return s;
  }
  // used by the library steppers' solve_impl: one compute_step of the library (solve, then the final forces and
  // their norm), with the stepper's parameters and preconditioner; fills its optim_result
  void run_library_solve(ssim_stepper_t kind, const ssim_solver_params* params, int precond,
                         mp::optim::optim_result<double>& result) {
    internal::check(ssim_set_stepper(ctx_, kind));
    ssim_solver_params p{};
    internal::check(ssim_get_solver_params(ctx_, &p));
    if (params) p = *params;
    if (precond >= 0) p.pcg_precond = precond;
    internal::check(ssim_set_solver_params(ctx_, &p));
    double r = 0;
    internal::check(ssim_compute_step(ctx_, &r));
    const ssim_step_stats st = last_stats();
    result.converged_ = st.converged != 0;
    result.iterations_ = st.newton_iterations;
    result.grad_norm_ = st.grad_norm;
    result.last_change_ = st.last_change;
    result.value_ = st.energy;
    residual_ = r;
  }

 private:
  template <typename A, typename = void>
  struct has_kind : std::false_type {};
  template <typename A>
  struct has_kind<A, std::void_t<decltype(A::kind)>> : std::true_type {};

  batched_vertex array(ssim_array_t which) {
    double* p = nullptr;
    internal::check(ssim_device_array(ctx_, which, &p));
    return {p, mp::make_shape(nv_, mp::holder<3>{})};
  }
  double param(int which) const {
    double v[6] = {0, 0, 0, 0, 0, 0};
    internal::check(ssim_get_physical_params(ctx_, &v[0], &v[1], &v[2], &v[3], &v[4], &v[5]));
    return v[which];
  }
  basic_time_step& chk(ssim_status_t st) {
    internal::check(st);
    return *this;
  }
  void rest_shape() {
    h_dminv_.resize(9 * static_cast<size_t>(nt_));
    h_vol_.resize(static_cast<size_t>(nt_));
    internal::check(ssim_get_rest_shape(ctx_, h_dminv_.data(), h_vol_.data()));
  }
  template <typename Export>
  sys_matrix_view export_csr(host_sys_matrix& m, Export fn) {
    long long nnz = 0;
    internal::check(ssim_csr_nnz(ctx_, &nnz));
    if (m.nnz() != nnz || m.rows() != total_dofs())
      m = host_sys_matrix(total_dofs(), total_dofs(), static_cast<index_t>(nnz), mp::sparse::sparse_property::symmetric);
    internal::check(fn(ctx_, m.outer_data(), m.inner_data(), m.values_data()));
    return m.view();
  }

  ssim_ctx* ctx_ = nullptr;
  index_t nv_ = 0, nt_ = 0;
  Scalar residual_ = 0;
  bool last_spsd_ = false;
  Blas blas_;
  ParImpl pf_;
  host_sys_matrix h_sys_, h_mass_, h_massf_;
  std::vector<double> h_dminv_, h_vol_, h_stiff_;
};

}  // namespace ssim::fem
