/*
 * ssim_hip.h -- C ABI of the MI355X-native FEM soft-body time step (libssim_hip.so).
 *
 * One context = one tetrahedral mesh with its state, living on the current HIP device.  Every entry point
 * returns an ssim_status_t; on failure ssim_last_error() names the cause.  No C++ exception crosses the ABI.
 * Arrays are plain host pointers with the reference's row-major layouts (vertices [nv,3] float64, cells
 * [nt,4] int32, vertex arrays [nv,3] flattened to 3·nv doubles, dof index 3v+c), except where an entry point
 * says "device".  All GPU work of a context runs on its stream (ssim_set_stream; default: the null stream).
 *
 * Each declaration cites the reference interface it replaces (paths relative to the reference tree,
 * Adversarr/ssim).
 */
#ifndef SSIM_HIP_H
#define SSIM_HIP_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ssim_ctx ssim_ctx;

typedef enum {
  SSIM_OK = 0,
  SSIM_ERR_INVALID_ARGUMENT = 1, /* the reference throws std::invalid_argument here */
  SSIM_ERR_HIP = 2,              /* a HIP runtime call failed (out of memory, no device, ...) */
  SSIM_ERR_STATE = 3,            /* call order violated, e.g. stepping before reset */
  SSIM_ERR_NUMERIC = 4           /* NaN/Inf in the state, or a solver breakdown */
} ssim_status_t;

/* Newton (time_step_solver_backward_euler), L-BFGS (time_step_solver_lbfgs), nonlinear CG (time_step_solver_ncg,
 * stepper/nonlinear_cg.hpp:9-37; it shares the L-BFGS iteration limit and line-search constants), forward Euler
 * (time_step_solver_forward_euler, finite_elements/time_step.hpp:717-736: next_deform += ½·dt²·forces, with
 * forces evaluated once at the predictor) */
typedef enum {
  SSIM_STEPPER_NEWTON = 0,
  SSIM_STEPPER_LBFGS = 1,
  SSIM_STEPPER_NCG = 2,
  SSIM_STEPPER_FORWARD_EULER = 3,
  SSIM_STEPPER_NCG_EXT_PREC = 4, /* NCG preconditioned by z = L·Lᵀ·g (time_step_solver_ncg_with_ext_prec,
                                    stepper/nonlinear_cg.hpp:41-97; L set by ssim_set_precond_matrix) */
  SSIM_STEPPER_LBFGS_PD = 5,      /* L-BFGS preconditioned by solves with the projective-dynamics matrix
                                    M + dt²κK built at reset (time_step_solver_pd, stepper/projective_dynamics.hpp:9-88;
                                    tet_lbfgs_pd, bindings/src/fem_host.cpp:223) */
  SSIM_STEPPER_LBFGS_H0 = 6       /* L-BFGS preconditioned by solves with the SPD Hessian at the start of each step
                                    (time_step_solver_h0, stepper/projective_dynamics.hpp:91-148) */
} ssim_stepper_t;

/* material model, the ElastModel template argument of basic_time_step (finite_elements/time_step.hpp:168-172):
 * stable Neo-Hookean (elast/stable_neohookean.hpp, the default) or linear elasticity (elast/linear.hpp) */
typedef enum { SSIM_MATERIAL_STABLE_NEOHOOKEAN = 0, SSIM_MATERIAL_LINEAR = 1 } ssim_material_t;

/* vertex-shaped arrays addressable through ssim_get_array / ssim_set_array / ssim_device_array
 * (accessors of finite_elements/time_step.hpp:275-311) */
typedef enum {
  SSIM_ARRAY_DEFORMATION = 0,   /* u at the current step         (deformation_)   */
  SSIM_ARRAY_PREV_DEFORM = 1,   /* u at the previous step        (prev_deform_)   */
  SSIM_ARRAY_INERTIA = 2,       /* predictor û                   (inertia_deform_)*/
  SSIM_ARRAY_NEXT_DEFORM = 3,   /* solver iterate / next u       (next_deform_)   */
  SSIM_ARRAY_VELOCITY = 4,      /*                               (velocity_)      */
  SSIM_ARRAY_EXT_ACCEL = 5,     /* external acceleration         (ext_accel_)     */
  SSIM_ARRAY_FORCES = 6,        /* gradient of the incremental potential (forces_) */
  SSIM_ARRAY_DBC_VALUES = 7,    /* Dirichlet values              (dbc_values_)    */
  SSIM_ARRAY_COUNT = 8
} ssim_array_t;

/* Solver parameters.  Reference: stepper/backward_euler.hpp:20-22 (Newton: tol = threshold·mass, Armijo 0.3)
 * and the benchmark configuration (≤10 Newton iterations, Jacobi-PCG rtol 1e-8). */
typedef struct {
  int max_newton;      /* default 10 */
  double pcg_rtol;     /* default 1e-8: stop when ‖r‖₂ ≤ rtol·‖b‖₂ */
  int pcg_maxit;       /* default 10000 */
  double ls_c;         /* Armijo constant, default 0.3 */
  double ls_beta;      /* backtracking factor, default 0.5 */
  int ls_max;          /* maximum trial steps per line search, default 20 */
  int lbfgs_m;         /* L-BFGS history, default 8 */
  int lbfgs_maxit;     /* default 1000 */
  double lbfgs_c;      /* Armijo constant for L-BFGS, default 1e-4 */
  int lbfgs_ls_max;    /* default 30 */
  int fp32_elements;   /* 1: element energy/stress/Hessian in FP32 (reductions and PCG stay FP64) */
  int newton_polish;   /* full Newton steps (α = 1, no line search, unprojected Hessian) after the Newton loop,
                          default 0 (the reference's algorithm); 1-2 take the solution to roundoff where the
                          Armijo test can no longer resolve energy decreases, as finite-difference checks of
                          gradients need.  A Hessian that is not positive definite there fails the PCG
                          (SSIM_ERR_NUMERIC) */
  int pcg_precond;     /* preconditioner of the CG solves with the system matrix: 0 = Jacobi (default; mathprim
                          diagonal_preconditioner, tests/fem_cu/main.cu:52-54), 1 = none (plain CG, as the
                          reference's Python bindings, bindings/src/fem_host.cpp:50-51).  Batches take 0 only. */
} ssim_solver_params;

/* Counters of the last compute_step (mathprim optim_result of stepper/backward_euler.hpp:28-32, plus
 * per-step work counts). */
typedef struct {
  int converged;
  int newton_iterations; /* Newton (or L-BFGS) iterations */
  int pcg_iterations;    /* total PCG iterations (SpMVs) */
  int energy_evals;      /* energy + gradient evaluations (line-search trials included) */
  int hessian_evals;
  int ls_failures;       /* line searches that exhausted ls_max trials */
  double grad_norm;      /* ‖forces‖₂ after the step */
  double energy;         /* incremental potential at the result */
  double last_change;    /* |E_k − E_{k−1}|: change of the incremental potential over the last accepted iteration
                            (optim_result::last_change_, stepper/backward_euler.hpp:31; 0 when no iteration ran) */
  int ls_rejected;       /* Newton: 1 when the last line search exhausted its trials and its smallest step raised
                            the energy; that step was not taken (u and forces stay at the last accepted iterate)
                            and the solve stopped unconverged */
  int reserved;
} ssim_step_stats;

const char* ssim_last_error(void);

/* basic_time_step(mesh, dt, E, ν, ρ) + reset(): finite_elements/time_step.hpp:260-265, :368-494.
 * Validates E ≥ 0, 0 < ν < 0.5, ρ ≥ 0, dt > 0 and every cell index. */
ssim_status_t ssim_create(ssim_ctx** out, const double* vertices, int nv, const int* cells, int nt,
                          double time_step, double youngs, double poisson, double density);
ssim_status_t ssim_destroy(ssim_ctx* ctx);
/* all later GPU work of ctx goes to this hipStream_t (passed as void*) */
ssim_status_t ssim_set_stream(ssim_ctx* ctx, void* hip_stream);

/* setters: finite_elements/time_step.hpp:316-347 (same validation; take effect at the next reset) */
ssim_status_t ssim_set_time_step(ssim_ctx* ctx, double dt);
ssim_status_t ssim_set_youngs(ssim_ctx* ctx, double youngs);
ssim_status_t ssim_set_poisson(ssim_ctx* ctx, double poisson);
ssim_status_t ssim_set_density(ssim_ctx* ctx, double density);
ssim_status_t ssim_set_damping(ssim_ctx* ctx, double damping);
/* set_threshold (time_step.hpp:343-347) / set_rtol (bindings/src/fem_host.cpp:146); like the reference, the
 * absolute tolerance rtol·mass is recomputed at the next reset */
ssim_status_t ssim_set_threshold(ssim_ctx* ctx, double rtol);
/* the getters time_step(), youngs(), poisson(), density(), damping() and grad_convergence_threshold()
 * (time_step.hpp:312, 349-353): the values last set; any output may be NULL */
ssim_status_t ssim_get_physical_params(const ssim_ctx* ctx, double* time_step, double* youngs, double* poisson,
                                       double* density, double* damping, double* threshold);
/* per-tet Lamé parameters (extension: the reference's per-call (λ, μ) overloads,
 * elast/basic_elast.hpp:187-194); NULL restores the uniform (E, ν) material */
ssim_status_t ssim_set_material_per_tet(ssim_ctx* ctx, const double* lambda, const double* mu);
ssim_status_t ssim_set_solver_params(ssim_ctx* ctx, const ssim_solver_params* params);
ssim_status_t ssim_get_solver_params(const ssim_ctx* ctx, ssim_solver_params* params);
/* Newton or L-BFGS (stepper/backward_euler.hpp), nonlinear CG (stepper/nonlinear_cg.hpp) or forward Euler
 * (finite_elements/time_step.hpp:717-736) */
ssim_status_t ssim_set_stepper(ssim_ctx* ctx, ssim_stepper_t stepper);
/* selects the ElastModel of basic_time_step<..., ElastModel> at run time; takes effect at the next evaluation */
ssim_status_t ssim_set_material_model(ssim_ctx* ctx, ssim_material_t model);
ssim_status_t ssim_get_material_model(const ssim_ctx* ctx, ssim_material_t* model);
/* Batched material evaluation on the device: the reference's vmap over elast::<model>::stress_op() and
 * hessian_op() (elast/basic_elast.hpp:144-219, tests/fem_intro/main.cpp:302-309), in 3D or 2D (the Ndim = 2
 * models, elast/stable_neohookean.hpp:96-107 and elast/linear.hpp).  F: n row-major dim×dim matrices; lambda,
 * mu: n Lamé pairs; outputs (each may be NULL) Ψ [n], P [n][dim²] and the Hessian [n][dim⁴] with
 * H[(a·dim+c)·dim² + (b·dim+d)] = ∂²Ψ/∂F_ac∂F_bd; make_spsd: the time step's SPD projection (eigenvalues of H
 * clamped from below at 1e-3, time_step.hpp:41-59, in either dimension).  All
 * arrays in device memory; launched on `stream` (NULL: the null stream), not synchronised. */
ssim_status_t ssim_material_eval(int dim, ssim_material_t model, long long n, const double* F, const double* lambda,
                                 const double* mu, int make_spsd, double* psi, double* P, double* H, void* stream);

/* P1 triangles in 2D: the per-element quantities of basic_time_step<..., PhysicalDim 2, TopologyDim 3>, which
 * the reference templates but never instantiates: the rest area |det Dm|/2 (finite_elements/rest_vol.hpp:33-50),
 * Dm⁻¹ and F = Ds·Dm⁻¹ (finite_elements/def_grad.hpp:53-127), the energy A·Ψ(F), the vertex gradient A·P·Dm⁻ᵀ
 * (stress_element_to_vertex, time_step.hpp:62-83), the 6×6 stiffness A·(∂F/∂x)ᵀ·H·(∂F/∂x) with H optionally
 * SPD-projected (eigenvalues clamped at 1e-3, time_step.hpp:41-59; ∂F/∂x: compute_pfpx_2,
 * elast/internal_physic.hpp:78-100), and the P1 consistent mass (times `density`) and Laplacian
 * (finite_elements/bilinears.hpp:65-91, 191-264).  X [nv][2] rest positions, u [nv][2] displacements,
 * cells [nt][3], lambda / mu [nt]; outputs, each may be NULL: area [nt], dminv [nt][2][2] and F [nt][2][2]
 * (row-major), energy [nt], grad [nt][3][2], stiff [nt][6][6] (dof 2·vertex + component), mass [nt][3][3],
 * laplacian [nt][3][3].  A triangle with a vertex index outside [0, nv) gets NaN outputs; a degenerate one
 * (det Dm = 0) gets area and mass 0 and non-finite Dm⁻¹, F, energy, gradient, stiffness and Laplacian (the
 * division by det Dm).  All arrays in device memory; launched on `stream` (NULL: the null stream), not
 * synchronised. */
ssim_status_t ssim_tri_eval(ssim_material_t model, int nv, int nt, const double* X, const int* cells, const double* u,
                            const double* lambda, const double* mu, double density, int make_spsd, double* area,
                            double* dminv, double* F, double* energy, double* grad, double* stiff, double* mass,
                            double* laplacian, void* stream);

/* mark_dirichlet_batched / mark_general_batched: bindings/src/fem_host.cpp:123-144 (targets [count,3]) */
ssim_status_t ssim_mark_dirichlet(ssim_ctx* ctx, const int* verts, int count, const double* targets);
ssim_status_t ssim_mark_general(ssim_ctx* ctx, const int* verts, int count);
/* add_ext_force_dof: finite_elements/time_step.hpp:550-557 (add_gravity = three calls, fem_host.cpp:148-152) */
ssim_status_t ssim_add_ext_force_dof(ssim_ctx* ctx, int dof_idx, double value);

/* reset(solver): recompute rest data, sparsity, assembly maps and thresholds (time_step.hpp:368-494) */
ssim_status_t ssim_reset(ssim_ctx* ctx);
/* prepare_step / compute_step / step_next: finite_elements/time_step.hpp:499-543 */
ssim_status_t ssim_prepare_step(ssim_ctx* ctx);
ssim_status_t ssim_compute_step(ssim_ctx* ctx, double* residual_out);
ssim_status_t ssim_step_next(ssim_ctx* ctx);
/* step(): prepare + compute + step_next (bindings/src/fem_host.cpp:102-111) */
ssim_status_t ssim_step(ssim_ctx* ctx, double* residual_out);
/* restore(): u_prev, u ← DBC values; v = (u − u_prev)/dt (time_step.hpp:560-569) */
ssim_status_t ssim_restore(ssim_ctx* ctx);
ssim_status_t ssim_get_step_stats(const ssim_ctx* ctx, ssim_step_stats* stats);

/* update_energy_and_gradients at next_deform: time_step.hpp:577-624 (forces_ ← g) */
ssim_status_t ssim_update_energy_and_gradients(ssim_ctx* ctx, double* energy_out);
/* update_hessian(make_spsd, filter_dbc) at next_deform: time_step.hpp:626-658 */
ssim_status_t ssim_update_hessian(ssim_ctx* ctx, int make_spsd, int filter_dbc);
/* ‖forces‖₂ < threshold·mass: time_step.hpp:660-664 */
ssim_status_t ssim_check_convergence(ssim_ctx* ctx, int* converged_out);

/* host copies of the vertex arrays (3·nv doubles) */
ssim_status_t ssim_get_array(ssim_ctx* ctx, ssim_array_t which, double* host_out);
ssim_status_t ssim_set_array(ssim_ctx* ctx, ssim_array_t which, const double* host_in);
/* device-to-device copies of a vertex array (3·nv doubles) from / into caller-owned device memory */
ssim_status_t ssim_get_array_device(ssim_ctx* ctx, ssim_array_t which, double* dst_device);
ssim_status_t ssim_set_array_device(ssim_ctx* ctx, ssim_array_t which, const double* src_device);
/* device pointer of a vertex array (3·nv doubles, owned by ctx, valid until destroy/reset) */
ssim_status_t ssim_device_array(ssim_ctx* ctx, ssim_array_t which, double** device_ptr);
/* Dirichlet flags per dof (dof_type_, 0 = general, 1 = dirichlet; boundary.hpp:8-11) */
ssim_status_t ssim_get_dof_type(const ssim_ctx* ctx, unsigned char* host_out);
/* device pointer of the Dirichlet flags (3·nv bytes, owned by ctx, valid until destroy): dof_type()
 * (time_step.hpp:279-280), which the reference's drivers write from device code (tests/fem_cu/main.cu:41-45).
 * Flags written there take effect like ssim_mark_dirichlet / ssim_mark_general: at once for the kernels, and
 * for the filtered mass matrix at the next ssim_reset. */
ssim_status_t ssim_device_dof_type(ssim_ctx* ctx, unsigned char** device_ptr);

ssim_status_t ssim_num_vertices(const ssim_ctx* ctx, int* nv);
ssim_status_t ssim_num_cells(const ssim_ctx* ctx, int* nt);
ssim_status_t ssim_total_mass(const ssim_ctx* ctx, double* mass);
ssim_status_t ssim_grad_threshold_abs(const ssim_ctx* ctx, double* tol);

/* scalar CSR with the full 3x3 block pattern, rows and columns ascending (the CSR of
 * finite_elements/global_composer.hpp:80-92 and time_step.hpp:456-461).  Any output may be NULL.
 * hessian(): the system matrix of the last update_hessian / Newton iteration (sysmatrix_, fem_host.cpp:169-171);
 * mass(): mass_matrix_ (fem_host.cpp:167). */
ssim_status_t ssim_csr_nnz(const ssim_ctx* ctx, long long* nnz);
/* tet_ncg_cuda_ext_ai::set_matrix (bindings/src/fem_cuda.cu:204): the preconditioner factor L of the
 * SSIM_STEPPER_NCG_EXT_PREC stepper, nnz values in the CSR order of ssim_export_hessian_csr (the system matrix's
 * pattern, nnz = ssim_csr_nnz); host array.  ssim_reset sets L to the identity (fem_cuda.cu:175-201). */
ssim_status_t ssim_set_precond_matrix(ssim_ctx* ctx, const double* values, long long nnz);
ssim_status_t ssim_export_hessian_csr(ssim_ctx* ctx, int* rowptr, int* colidx, double* values);
/* the projective-dynamics matrix of SSIM_STEPPER_LBFGS_PD (time_step_solver_pd::pd_matrix_,
 * stepper/projective_dynamics.hpp:51-87) in the same CSR layout; built at reset or on first use */
ssim_status_t ssim_export_pd_csr(ssim_ctx* ctx, int* rowptr, int* colidx, double* values);
ssim_status_t ssim_export_mass_csr(ssim_ctx* ctx, int* rowptr, int* colidx, double* values);
/* mass_matrix_filtered() (time_step.hpp:310-311, built at :460-466): the mass matrix with the Dirichlet rows and
 * columns replaced by the identity (boundary.hpp:53-67), for the current dof flags; same CSR layout */
ssim_status_t ssim_export_mass_filtered_csr(ssim_ctx* ctx, int* rowptr, int* colidx, double* values);
/* dminv() and rest_vol() (time_step.hpp:296-301): Dm⁻¹ per tet, [nt,3,3] row-major with Dm = [X1−X0, X2−X0,
 * X3−X0] (def_grad.hpp:56-91), and V = |det Dm|/6 per tet (rest_vol.hpp:34-54); host arrays, either may be NULL;
 * computed by ssim_reset */
ssim_status_t ssim_get_rest_shape(ssim_ctx* ctx, double* dminv, double* rest_vol);
/* mesh() (time_step.hpp:275-276): device pointers to the rest positions [nv,3] (double) and the cells [nt,4] (int),
 * owned by ctx and valid until ssim_destroy; either output may be NULL */
ssim_status_t ssim_device_mesh(const ssim_ctx* ctx, const double** vertices, const int** cells);

/* Element-level outputs at deformation u (host, 3·nv doubles): Ψ [nt], P [nt,3,3], H [nt,9,9] (SPD-projected
 * with floor 1e-3 when make_spsd; wrapped_pfpx, time_step.hpp:41-59) and the local stiffness blocks
 * K [nt,12,12] = pfpxᵀ·H·pfpx.  Outputs may be NULL.  stable_neohookean.hpp:23-91. */
ssim_status_t ssim_eval_elements(ssim_ctx* ctx, const double* u, int make_spsd, double* psi, double* P,
                                 double* H, double* K);
/* Jacobi-PCG on the current system matrix: solve H x = b (host arrays, 3·nv); mathprim cg +
 * diagonal_preconditioner (tests/fem_cu/main.cu:52-54) */
ssim_status_t ssim_pcg_solve(ssim_ctx* ctx, const double* b, double* x, double rtol, int maxit, int* iters_out);
/* y = H x with the current system matrix (host arrays) */
ssim_status_t ssim_spmv(ssim_ctx* ctx, const double* x, double* y);
/* ssim_pcg_solve on DEVICE vectors b and x (3·nv doubles each) that allocates nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (HIP graph-capture setups, serving loops, harnesses
 * that forbid hipMalloc).  Its work vectors, partial sums and PCG scalars live in a static 64 MiB device array
 * per GPU (up to about 2.0M dofs; larger meshes return SSIM_ERR_INVALID_ARGUMENT).  Same kernels and results as
 * ssim_pcg_solve.  Calls that overlap on one GPU share that array and race. */
ssim_status_t ssim_pcg_solve_noalloc(ssim_ctx* ctx, const double* b, double* x, double rtol, int maxit,
                                     int* iters_out);

/* ---- adjoint of one implicit step (no reference counterpart: the reference is forward-only) ----
 * At the converged u* = next_deform of the last compute_step, the step solves g(u; û, λ, μ) = 0 on the free dofs
 * with g = M(u − û) + dt²·Σ_e V_e f_e(u; λ_e, μ_e) (update_energy_and_gradients, time_step.hpp:577-624).  For
 * a loss with ∂L/∂u* = grad_u (Dirichlet entries ignored) this solves H l = grad_u with the unprojected,
 * Dirichlet-filtered Hessian (update_hessian(make_spsd = false), time_step.hpp:626-658) by Jacobi-PCG (the
 * solver parameters' rtol / maxit) and writes
 *   grad_uhat   = ∂L/∂û = M l (3·nv; zero on Dirichlet dofs),
 *   grad_lambda = ∂L/∂λ_e, grad_mu = ∂L/∂μ_e per tet (nt each; may be NULL),
 *   adjoint_out = l (3·nv; may be NULL).
 * All arrays are DEVICE pointers.  PCG breakdown (H not positive definite) returns SSIM_ERR_NUMERIC. */
ssim_status_t ssim_adjoint_step(ssim_ctx* ctx, const double* grad_u, double* grad_uhat, double* grad_lambda,
                                double* grad_mu, double* adjoint_out, int* pcg_iters_out);

/* ---- instrumentation (the reference times whole steps with std::chrono, bindings/src/fem_host.cpp:102-111;
 *      here kernels are timed on the GPU with hipEvents) ---- */
typedef enum {
  SSIM_KERNEL_ELEM_GRAD = 0,   /* per-tet energy + gradient contributions            */
  SSIM_KERNEL_GRAD_REDUCE = 1, /* per-vertex segmented reduce of the gradient        */
  SSIM_KERNEL_ELEM_HESS = 2,   /* per-tet SNH Hessian, SPD projection, 3x3 blocks    */
  SSIM_KERNEL_HESS_REDUCE = 3, /* per-block segmented reduce into the blocked ELL    */
  SSIM_KERNEL_PCG_SPMV = 4,    /* p = z + βp fused with q = A p and p·q (PCG kernel A) */
  SSIM_KERNEL_PCG_UPDATE = 5,  /* x, r, z updates with r·z and r·r (PCG kernel B)    */
  SSIM_KERNEL_PCG_P = 6,       /* reserved (the p update is fused into kernel A)     */
  SSIM_KERNEL_COUNT = 7
} ssim_kernel_t;

/* Enable (1) or disable (0) sampled kernel timing and clear its accumulators.  When enabled, one launch of each
 * kernel per call site and per PCG chunk is timed by two hipEvents on the context's stream; for the single-mesh
 * element kernels (tile-reduced assemblies) and the PCG kernels the runtime stamps them at the kernel's own start
 * and end, for the others they bracket the launch. */
ssim_status_t ssim_set_kernel_timing(ssim_ctx* ctx, int enabled);
/* average duration in milliseconds and number of sampled launches of one kernel */
ssim_status_t ssim_get_kernel_timing(const ssim_ctx* ctx, ssim_kernel_t kernel, double* avg_ms, long long* samples);

/* sizes of the device layout (for byte accounting) */
typedef struct {
  int nv, nt;
  long long nonzero_blocks;  /* 3x3 blocks of the sparsity pattern (nnz / 9) */
  long long padded_slots;    /* stored (upper, row ≤ column) blocked-ELL slots including padding */
  long long lower_slots;     /* lower-part reference slots including padding */
  int upper_blocks;          /* blocks with row ≤ column */
  int slices;                /* 64-row slices */
  int fixed_layout;          /* 1: every row padded to 7 upper and 7 lower off-diagonal blocks, diagonal blocks
                                stored as their 6-value upper triangle, 16-bit column offsets (32 B of indices per
                                row); 0: the general layout (32-bit columns and references, 9-value blocks) */
  int reserved;
} ssim_layout_info;
ssim_status_t ssim_get_layout_info(const ssim_ctx* ctx, ssim_layout_info* info);

/* ---- launch tuning (no reference counterpart: the reference's launch shapes live inside mathprim) ---- */
typedef enum {
  SSIM_TUNE_XCD_REMAP = 0,     /* 1 (default): each of the 8 XCDs takes a contiguous range of SpMV row slices and of tets */
  SSIM_TUNE_PCG_CHUNK_MAX = 1, /* PCG iterations enqueued per done-flag poll (default 64) */
  SSIM_TUNE_SPMV_LANES = 2,    /* SpMV variant: 0 = one lane per row, fully unrolled over the fixed layout (the
                                  only variant, and the one set, when the mesh has it: ssim_layout_info), 1 = one
                                  lane per row with loops (default otherwise), 4 = four lanes per row */
  SSIM_TUNE_PCG_GRAPH = 3,     /* 1: replay PCG chunks from captured hipGraphs; 0 (default): launch each kernel
                                  (measured the same speed for one 786k-tet mesh and for 256 batched meshes: the gap
                                  between dependent kernels is not the host's launch cost) */
  SSIM_TUNE_COMPACT_PCT = 4,   /* batches: compact the PCG grids to the meshes still iterating once those hold at
                                  most this percentage of the current grid (default 70; 0 disables) */
  SSIM_TUNE_LBFGS_GRAM = 5,    /* L-BFGS with the diagonal H0 and m ≤ 8: 1 (default) runs the two-loop recursion on
                                  Gram matrices of the stored pairs (one fused dot-product pass and one fused
                                  direction pass per iteration), 0 runs it vector by vector */
  SSIM_TUNE_GRAD_TILE = 6,     /* tets per tile of the tile-reduced energy + gradient: 64, 128 or 256 (default);
                                  the tile maps are rebuilt on change.  Batches keep 256 (their per-mesh tet ranges
                                  are padded to 256-tet workgroups) */
  SSIM_TUNE_HESS_TILE = 7,     /* 1 (default): tile-reduced SPD Hessian assembly (one partial per 256-tet tile and
                                  upper block); 0: every tet writes its 10 blocks to contribution slots that a
                                  per-block reduce reads back */
  SSIM_TUNE_REST_FROM_X = 8    /* 1 (default): the energy + gradient and tile-reduced Hessian kernels form Dm⁻¹ and
                                  the rest volume from the gathered rest positions (the rest-shape kernel's
                                  arithmetic, bitwise the stored values); 0: they load the stored ones (80 B/tet) */
} ssim_tune_t;
ssim_status_t ssim_set_tuning(ssim_ctx* ctx, ssim_tune_t key, int value);
ssim_status_t ssim_get_tuning(const ssim_ctx* ctx, ssim_tune_t key, int* value);

/* ---- batches: many independent meshes stepped by one set of launches ----
 * No reference counterpart: the reference steps one mesh per finite_elements/time_step.hpp object.  Each mesh
 * of a batch follows that object's Newton step exactly (backward_euler.hpp:15-33, same threshold, line search
 * and PCG limits, per-mesh threshold·mass tolerance); the batch advances the meshes' Newton iterations in lock
 * step and every PCG iteration or line-search trial runs as one launch pair / evaluation over all meshes that
 * still take part.  Meshes have per-tet Lamé parameters; density and time step are shared.  A mesh index of −1
 * in the array calls means "all meshes, concatenated in batch order". */
typedef struct ssim_batch ssim_batch;
/* vertices: Σnv[m] rows of 3 doubles; cells: Σnt[m] rows of 4 mesh-local vertex indices; lambda, mu: Σnt[m]
 * (replaces one time_step constructor + reset() per mesh, time_step.hpp:260-265, 368-494) */
ssim_status_t ssim_batch_create(ssim_batch** out, int nmesh, const int* nv, const int* nt, const double* vertices,
                                const int* cells, const double* lambda, const double* mu, double density,
                                double time_step);
ssim_status_t ssim_batch_destroy(ssim_batch* batch);
ssim_status_t ssim_batch_set_stream(ssim_batch* batch, void* hip_stream);
ssim_status_t ssim_batch_num_meshes(const ssim_batch* batch, int* nmesh);
ssim_status_t ssim_batch_set_threshold(ssim_batch* batch, double rtol);  /* set_threshold, per mesh ×mass */
ssim_status_t ssim_batch_set_solver_params(ssim_batch* batch, const ssim_solver_params* params);
ssim_status_t ssim_batch_get_solver_params(const ssim_batch* batch, ssim_solver_params* params);
/* mark_dirichlet_batched / mark_general_batched of one mesh (bindings/src/fem_host.cpp:123-144), mesh-local
 * indices */
ssim_status_t ssim_batch_mark_dirichlet(ssim_batch* batch, int mesh, const int* verts, int count,
                                        const double* targets);
ssim_status_t ssim_batch_mark_general(ssim_batch* batch, int mesh, const int* verts, int count);
ssim_status_t ssim_batch_reset(ssim_batch* batch);
/* prepare_step / compute_step / step_next / step of every mesh (time_step.hpp:499-569) */
ssim_status_t ssim_batch_prepare_step(ssim_batch* batch);
ssim_status_t ssim_batch_compute_step(ssim_batch* batch);
ssim_status_t ssim_batch_step_next(ssim_batch* batch);
ssim_status_t ssim_batch_step(ssim_batch* batch);
/* statistics of the last compute_step: per mesh, or (mesh = −1) for the batch: lock-step Newton iterations,
 * PCG iterations summed over Newton iterations of the slowest mesh, batched evaluations, all converged */
ssim_status_t ssim_batch_get_step_stats(const ssim_batch* batch, int mesh, ssim_step_stats* stats);
ssim_status_t ssim_batch_get_array(ssim_batch* batch, ssim_array_t which, int mesh, double* host_out);
ssim_status_t ssim_batch_set_array(ssim_batch* batch, ssim_array_t which, int mesh, const double* host_in);
ssim_status_t ssim_batch_total_mass(const ssim_batch* batch, int mesh, double* mass);
/* device-to-device copies of a vertex array of all meshes, compact layout (Σ 3·nv[m] doubles) */
ssim_status_t ssim_batch_get_array_device(ssim_batch* batch, ssim_array_t which, double* dst_device);
ssim_status_t ssim_batch_set_array_device(ssim_batch* batch, ssim_array_t which, const double* src_device);
/* replace every mesh's per-tet Lamé parameters (host arrays of Σnt[m]); takes effect at once */
ssim_status_t ssim_batch_set_material(ssim_batch* batch, const double* lambda, const double* mu);
/* material model of every mesh of the batch (see ssim_set_material_model) */
ssim_status_t ssim_batch_set_material_model(ssim_batch* batch, ssim_material_t model);
/* ssim_adjoint_step for every mesh of the batch; device arrays in the compact layout (grad_u, grad_uhat,
 * adjoint_out: Σ 3·nv[m]; grad_lambda, grad_mu: Σ nt[m]); pcg_iters_out: host array of nmesh (may be NULL) */
ssim_status_t ssim_batch_adjoint_step(ssim_batch* batch, const double* grad_u, double* grad_uhat, double* grad_lambda,
                                      double* grad_mu, double* adjoint_out, int* pcg_iters_out);
/* layout of the concatenated, padded mesh */
ssim_status_t ssim_batch_get_layout_info(const ssim_batch* batch, ssim_layout_info* info);
/* launch tuning (same keys; SpMV variant 0 or 1) and sampled kernel timing of the batch's launches */
ssim_status_t ssim_batch_set_tuning(ssim_batch* batch, ssim_tune_t key, int value);
ssim_status_t ssim_batch_get_tuning(const ssim_batch* batch, ssim_tune_t key, int* value);
ssim_status_t ssim_batch_set_kernel_timing(ssim_batch* batch, int enabled);
ssim_status_t ssim_batch_get_kernel_timing(const ssim_batch* batch, ssim_kernel_t kernel, double* avg_ms,
                                           long long* samples);

#ifdef __cplusplus
}
#endif
#endif
