/*
 * ssim_oracle.h -- CPU FP64 restatement of the reference FEM soft-body time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (ssim_amd/, include/) may link, load or call this
 * code.  The tests compare the HIP path against it; tests/ and bench.py's cpu_baseline
 * leg load it through ctypes (oracle/oracle.py).
 *
 * Every routine restates one piece of the reference (citations are relative to the reference tree,
 * Adversarr/ssim).  The reference cannot be built here (its mathprim dependency is not vendored and is
 * fetched from the network by get-mathprim.sh), so the solver routines (PCG, Newton line search, L-BFGS)
 * restate the textbook algorithms with the parameters the benchmark configurations name; see the header of
 * ssim_oracle.c for which functions are pinned by which check.
 */
#ifndef SSIM_ORACLE_H
#define SSIM_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* ---- element level (row-major 3x3 matrices, 9x9 Hessian index 3a+c <-> F[a][c]) ---- */
double orc_mu(double E, double nu);
double orc_lambda(double E, double nu);
double orc_snh_energy(const double F[9], double lam, double mu);
double orc_snh_stress(const double F[9], double lam, double mu, double P[9]);
void orc_snh_hessian(const double F[9], double lam, double mu, double H[81]);
int orc_sym_eig(int n, const double* A, double* w, double* Q); /* Q columns are eigenvectors */
void orc_spd_project9(const double H[81], double floor_, double Hp[81]);
void orc_dminv(const double x0[3], const double x1[3], const double x2[3], const double x3[3], double B[9],
               double* V);
void orc_def_grad(const double x0[3], const double x1[3], const double x2[3], const double x3[3],
                  const double B[9], double F[9]);
void orc_pfpx(const double B[9], double pfpx[108]);
void orc_local_force(const double P[9], const double B[9], double f[12]);
void orc_local_stiffness(const double H[81], const double B[9], double K[144]);

/* ---- problem level: mirrors ssim::fem::basic_time_step (time_step.hpp:171-715) ---- */
typedef struct orc_ts orc_ts;

orc_ts* orc_ts_create(const double* X, int nv, const int* cells, int nt, double dt, const double* lam,
                      const double* mu, double rho);
void orc_ts_destroy(orc_ts* ts);
int orc_ts_nnz(const orc_ts* ts);
void orc_ts_csr_pattern(const orc_ts* ts, int* rowptr, int* colidx);
void orc_ts_mass_values(const orc_ts* ts, double* vals);
double orc_ts_total_mass(const orc_ts* ts);
void orc_ts_rest_data(const orc_ts* ts, double* B, double* V);

void orc_ts_set_dirichlet(orc_ts* ts, const unsigned char* dof_type, const double* dbc_values);
void orc_ts_set_ext_accel(orc_ts* ts, const double* accel);
void orc_ts_set_threshold(orc_ts* ts, double rtol);
void orc_ts_set_state(orc_ts* ts, const double* u, const double* u_prev, const double* vel);
void orc_ts_get_state(const orc_ts* ts, double* u, double* u_prev, double* vel);
void orc_ts_set_vectors(orc_ts* ts, const double* u_next, const double* u_hat);
void orc_ts_get_vectors(const orc_ts* ts, double* u_next, double* u_hat, double* forces);
void orc_ts_set_solver(orc_ts* ts, int max_newton, double pcg_rtol, int pcg_maxit, double ls_c,
                       double ls_beta, int ls_max);
/* full Newton steps (α = 1) after the Newton loop of compute_step (default 0) */
void orc_ts_set_polish(orc_ts* ts, int polish);
/* material model: 0 stable Neo-Hookean (default), 1 linear elasticity */
void orc_ts_set_model(orc_ts* ts, int model);
/* NCG preconditioner factor L (z = L·Lᵀ·g) in the Hessian's CSR order; NULL removes it */
void orc_ts_set_precond(orc_ts* ts, const double* values);
/* projective dynamics: A = M + dt²κK (Dirichlet-filtered) and L-BFGS preconditioned by solves with it */
void orc_ts_build_pd(orc_ts* ts, double kappa);
void orc_ts_pd_values(const orc_ts* ts, double* vals);
double orc_ts_compute_step_lbfgs_pd(orc_ts* ts, double* stats);
/* L-BFGS preconditioned by solves with the SPD Hessian at the start of the step */
double orc_ts_compute_step_lbfgs_h0(orc_ts* ts, double* stats);
double orc_lin_energy(const double F[9], double lam, double mu);
double orc_lin_stress(const double F[9], double lam, double mu, double P[9]);
void orc_lin_hessian(double lam, double mu, double H[81]);
/* 2D models (F row-major 2x2, Hessian 4x4) */
double orc_snh2_energy(const double F[4], double lam, double mu);
double orc_snh2_stress(const double F[4], double lam, double mu, double P[4]);
void orc_snh2_hessian(const double F[4], double lam, double mu, double H[16]);
double orc_lin2_energy(const double F[4], double lam, double mu);
double orc_lin2_stress(const double F[4], double lam, double mu, double P[4]);
void orc_lin2_hessian(double lam, double mu, double H[16]);
/* P1 triangles in 2D, per element (model 0 stable Neo-Hookean, 1 linear): area [nt], Dm⁻¹ and F [nt][4]
 * (row-major), A·Ψ [nt], vertex gradient [nt][6], stiffness [nt][36] (dof 2a+c), ρ-scaled mass and Laplacian
 * [nt][9]; outputs may be NULL; a triangle with an index outside [0, nv) gets NaN */
void orc_tri_eval(int model, int nv, int nt, const double* X, const int* cells, const double* u, const double* lam,
                  const double* mu, double rho, int make_spsd, double* area, double* dminv, double* F,
                  double* energy, double* grad, double* stiff, double* mass, double* lap);
/* forward Euler: u_next = û + ½dt²·g(û) (time_step.hpp:717-736) */
double orc_ts_compute_step_fe(orc_ts* ts, double* stats);
void orc_ts_set_lbfgs(orc_ts* ts, int m, int max_iter, double ls_c, double ls_beta, int ls_max);

/* element outputs at deformation u (any NULL output is skipped) */
void orc_ts_elements(const orc_ts* ts, const double* u, double* psi, double* P, double* H, double* Hp,
                     double* K, int make_spsd);
double orc_ts_energy_grad(orc_ts* ts, const double* u, double* g);
void orc_ts_hessian(orc_ts* ts, const double* u, int make_spsd, int filter_dbc, double* vals);

int orc_pcg(int n, const int* rowptr, const int* colidx, const double* vals, const double* b, double* x,
            double rtol, int maxit, double* out_rel_resid);

void orc_ts_prepare_step(orc_ts* ts);
/* stats: [0]=newton iterations, [1]=pcg iterations total, [2]=energy evaluations, [3]=converged flag,
 *        [4]=final grad norm, [5]=line-search failures */
double orc_ts_compute_step(orc_ts* ts, double* stats);
double orc_ts_compute_step_lbfgs(orc_ts* ts, double* stats);
/* nonlinear CG (Dai-Kou DK+ β, backtracking Armijo with the L-BFGS constants); stats as compute_step */
double orc_ts_compute_step_ncg(orc_ts* ts, double* stats);
void orc_ts_step_next(orc_ts* ts);
/* velocity damping in [0, 1) (returns 1 and keeps the old value otherwise) */
int orc_ts_set_damping(orc_ts* ts, double damping);
void orc_ts_rest_hessian_diag(orc_ts* ts, double* diag);

/* timing helpers for the CPU baseline: one PCG iteration-equivalent (SpMV + vector work) etc. */
void orc_ts_spmv(const orc_ts* ts, const double* vals, const double* x, double* y);
int orc_num_threads(void);
void orc_set_num_threads(int n); /* thread count of the following parallel loops (OpenMP builds) */

#ifdef __cplusplus
}
#endif
#endif
