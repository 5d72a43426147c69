/*
 * ssim_oracle.c -- CPU FP64 restatement of the reference FEM soft-body implicit time step.
 *
 * TEST INFRASTRUCTURE ONLY: what the HIP path is tested against, never part of the product.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it (through oracle/oracle.py).
 *
 * Citations are relative to the reference tree (Adversarr/ssim).  Conventions follow the reference:
 * row-major 3x3 matrices; F[a][c] = Σ_k Ds[a][k] B[k][c]; the 9x9 F-space Hessian is indexed 3a+c;
 * the 12x12 local stiffness is indexed 3v+b (vertex v, component b); dof index of vertex v is 3v+b.
 *
 * What pins each part (tests/test_oracle.py):
 *   orc_snh_*            closed forms Ψ(I)=0, P(I)=0 (tests/fem_intro/main.cpp:100-106), central FD of Ψ and P
 *                        at the reference test point (tests/fem_intro/main.cpp:109), torch FP64 autograd
 *                        fixtures (tests/golden/snh_autograd.json).
 *   orc_sym_eig /        A Q = Q diag(w), QᵀQ = I, numpy.linalg.eigh fixtures, invariants of the clamp
 *   orc_spd_project9     (H⁺ = H when λ_min ≥ floor; eig(H⁺) = max(eig H, floor)).
 *   orc_dminv, def_grad  F = I at u = 0; ΣV = box volume (tests/fem_intro/main.cpp:311-333); affine u gives
 *                        the affine map; B·Dm = I.
 *   mass, assembly       Σ M = 3ρV_total and the per-tet closed form (tests/fem_local_global/main.cpp:166-184);
 *                        a dense brute-force assembly on tiny meshes (tests/fem_local_global/main.cpp:85-164).
 *   energy_grad, hessian central FD of E for g and of g for the unprojected H on `tetra` and a 5-tet
 *                        unit box (tests/fem_time_step_diff/main.cpp:19-172).
 *   orc_pcg              dense numpy solve on small SPD systems; iteration count bounded by n in exact math.
 *   compute_step         the Newton iterate is a stationary point: ‖g‖ below the stopping tolerance, and equals
 *                        scipy.optimize.minimize on a tiny mesh; energy decreases monotonically per iteration.
 *   compute_step_lbfgs   the same stationarity check and agreement with the Newton minimiser.
 *   orc_tri_eval         (tests/test_oracle_tri.py) area / Dm⁻¹ against numpy, affine F, closed-form mass, the
 *                        cotangent Laplacian, central FD of the energy and gradient, numpy.linalg.eigh for the
 *                        clamp, rotation invariance.
 *   prepare_step,        (tests/test_oracle_stepping.py) free fall from rest: u_n = ½dt²g·n(n+1), v_n = n·dt·g and
 *   step_next, damping   the damped closed form; a clamped, damped linear bar settles at K_ff u_f = M_ff g_f with
 *                        K and M assembled independently (pins the dt² scaling of the elastic energy).
 *   snh2_delta           (tests/test_oracle_2d.py) the property the reference's quadratic encodes: with the
 *                        stress-free α, the Hessian at F = 0 has zero curvature along pure shear
 *                        (μ(1 − 1/δ) + λα = 0, which times δ(2+δ)/λ is the quadratic), checked on the oracle's
 *                        Hessian and on the ε² decay of P(εS)/ε.
 * Parity unpinned:
 *   - the solver trajectories of mathprim (the reference's external solver library, not vendored): iteration
 *     order, stopping norm and line-search constants follow the benchmark configuration.
 */
#include "ssim_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------------------------------ */
/* Lamé parameters: elast/basic_elast.hpp:11-20                                                      */
double orc_mu(double E, double nu) { return E / (2 * (1 + nu)); }
double orc_lambda(double E, double nu) { return E * nu / ((1 + nu) * (1 - 2 * nu)); }

static double det3(const double A[9]) {
  return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
         A[2] * (A[3] * A[7] - A[4] * A[6]);
}

/* cofactor matrix rows: f1×f2, f2×f0, f0×f1 for rows f_k of F.
 * elast/internal_physic.hpp:40-48 (partial_determinant on the column-major map = Fᵀ, so its columns are
 * the rows of cof F). */
static void cofactor(const double F[9], double C[9]) {
  const double *f0 = F, *f1 = F + 3, *f2 = F + 6;
  C[0] = f1[1] * f2[2] - f1[2] * f2[1];
  C[1] = f1[2] * f2[0] - f1[0] * f2[2];
  C[2] = f1[0] * f2[1] - f1[1] * f2[0];
  C[3] = f2[1] * f0[2] - f2[2] * f0[1];
  C[4] = f2[2] * f0[0] - f2[0] * f0[2];
  C[5] = f2[0] * f0[1] - f2[1] * f0[0];
  C[6] = f0[1] * f1[2] - f0[2] * f1[1];
  C[7] = f0[2] * f1[0] - f0[0] * f1[2];
  C[8] = f0[0] * f1[1] - f0[1] * f1[0];
}

/* Stable Neo-Hookean, 3D, δ = 1: elast/stable_neohookean.hpp:23-36 (energy), :101-103 (δ) */
double orc_snh_energy(const double F[9], double lam, double mu) {
  const double del = 1.0;
  double Ic = 0;
  for (int i = 0; i < 9; ++i) Ic += F[i] * F[i];
  const double alpha = (1 - 1 / (3 + del)) * mu / lam + 1;
  const double Jm = det3(F) - alpha;
  const double zero_energy = 0.5 * (lam * (1 - alpha) * (1 - alpha) - mu * log(3 + del));
  return 0.5 * (mu * (Ic - 3) + lam * Jm * Jm - mu * log(Ic + del)) - zero_energy;
}

/* P = λ(J−α) cof F + μ(1 − 1/(Ic+δ)) F: elast/stable_neohookean.hpp:38-62 */
double orc_snh_stress(const double F[9], double lam, double mu, double P[9]) {
  const double del = 1.0;
  double Ic = 0;
  for (int i = 0; i < 9; ++i) Ic += F[i] * F[i];
  const double alpha = (1 - 1 / (3 + del)) * mu / lam + 1;
  const double Jm = det3(F) - alpha;
  double C[9];
  cofactor(F, C);
  const double s1 = lam * Jm, s2 = mu * (1.0 - 1.0 / (Ic + del));
  for (int i = 0; i < 9; ++i) P[i] = s1 * C[i] + s2 * F[i];
  const double zero_energy = 0.5 * (lam * (1 - alpha) * (1 - alpha) - mu * log(3 + del));
  return 0.5 * (mu * (Ic - 3) + lam * Jm * Jm - mu * log(Ic + del)) - zero_energy;
}

/* x̂(v) = [[0,−v2,v1],[v2,0,−v0],[−v1,v0,0]]: elast/internal_physic.hpp:58-63 */
static void add_xhat_block(double H[81], int ra, int cb, const double v[3], double s) {
  /* adds s·x̂(v) into the 3x3 block (rows 3ra.., cols 3cb..) */
  const double xh[9] = {0, -v[2], v[1], v[2], 0, -v[0], -v[1], v[0], 0};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) H[(3 * ra + i) * 9 + 3 * cb + j] += s * xh[i * 3 + j];
}

/* H = μ(1−1/(Ic+δ)) I + 2μ/(Ic+δ)² f fᵀ + λ g gᵀ + λ(J−α) H_J
 * elast/stable_neohookean.hpp:64-91, add_HJ elast/internal_physic.hpp:65-76 */
void orc_snh_hessian(const double F[9], double lam, double mu, double H[81]) {
  const double del = 1.0;
  double I2 = 0;
  for (int i = 0; i < 9; ++i) I2 += F[i] * F[i];
  const double I3 = det3(F);
  const double alpha = (1 - 1 / (3 + del)) * mu / lam + 1;
  const double scale = lam * (I3 - alpha);
  double g[9];
  cofactor(F, g);
  const double d = mu * (1 - 1 / (I2 + del));
  const double ff = mu * (2.0 / (I2 + del) / (I2 + del));
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) H[i * 9 + j] = (i == j ? d : 0.0) + ff * F[i] * F[j] + lam * g[i] * g[j];
  const double *f0 = F, *f1 = F + 3, *f2 = F + 6;
  add_xhat_block(H, 1, 0, f2, scale);
  add_xhat_block(H, 0, 1, f2, -scale);
  add_xhat_block(H, 2, 0, f1, -scale);
  add_xhat_block(H, 0, 2, f1, scale);
  add_xhat_block(H, 2, 1, f0, scale);
  add_xhat_block(H, 1, 2, f0, -scale);
}

/* Cyclic Jacobi eigenvalue algorithm for a symmetric n×n matrix (Golub & Van Loan, Matrix Computations,
 * 4th ed., Alg. 8.5.3, with Rutishauser's rotation formulas).  Stands in for Eigen's
 * SelfAdjointEigenSolver used at finite_elements/time_step.hpp:51.  Returns the number of sweeps. */
int orc_sym_eig(int n, const double* A_in, double* w, double* Q) {
  double A[81];
  for (int i = 0; i < n * n; ++i) A[i] = A_in[i];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) Q[i * n + j] = (i == j);
  double scale = 0;
  for (int i = 0; i < n * n; ++i) scale += A[i] * A[i];
  int sweep = 0;
  for (; sweep < 100; ++sweep) {
    double off = 0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += A[p * n + q] * A[p * n + q];
    if (off <= 1e-36 * scale || off == 0.0) break;
    for (int p = 0; p < n; ++p) {
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[p * n + q];
        if (apq == 0.0) continue;
        const double app = A[p * n + p], aqq = A[q * n + q];
        const double theta = (aqq - app) / (2 * apq);
        double t;
        if (fabs(theta) > 1e150)
          t = 0.5 / theta;
        else
          t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
        const double c = 1 / sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; ++k) {
          if (k == p || k == q) continue;
          const double akp = A[k * n + p], akq = A[k * n + q];
          const double nkp = c * akp - s * akq, nkq = s * akp + c * akq;
          A[k * n + p] = A[p * n + k] = nkp;
          A[k * n + q] = A[q * n + k] = nkq;
        }
        A[p * n + p] = app - t * apq;
        A[q * n + q] = aqq + t * apq;
        A[p * n + q] = A[q * n + p] = 0.0;
        for (int k = 0; k < n; ++k) {
          const double vkp = Q[k * n + p], vkq = Q[k * n + q];
          Q[k * n + p] = c * vkp - s * vkq;
          Q[k * n + q] = s * vkp + c * vkq;
        }
      }
    }
  }
  for (int i = 0; i < n; ++i) w[i] = A[i * n + i];
  return sweep;
}

/* H⁺ = Q diag(max(λ_i, floor)) Qᵀ: finite_elements/time_step.hpp:50-56 (floor 1e-3, before V and dt²) */
void orc_spd_project9(const double H[81], double floor_, double Hp[81]) {
  double w[9], Q[81];
  orc_sym_eig(9, H, w, Q);
  for (int i = 0; i < 9; ++i) w[i] = w[i] > floor_ ? w[i] : floor_;
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) {
      double s = 0;
      for (int k = 0; k < 9; ++k) s += Q[i * 9 + k] * w[k] * Q[j * 9 + k];
      Hp[i * 9 + j] = s;
    }
}

/* Linear elasticity (elast/linear.hpp:24-69, strain internal_physic.hpp:16-19): ε = ½(F + Fᵀ) − I,
 * Ψ = μ‖ε‖² + ½λ·tr(ε)², P = 2μ·ε + λ·tr(ε)·I, and the constant Hessian H = μ(I + S) + λ·vec(I)vec(I)ᵀ, S the
 * transposition of vec(F) (entry (3i+j, 3j+i) gets μ, entry (3i+i, 3j+j) gets λ, linear.hpp:55-68). */
static void lin_strain(const double F[9], double eps[9], double* tr) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) eps[3 * i + j] = 0.5 * (F[3 * i + j] + F[3 * j + i]) - (i == j ? 1.0 : 0.0);
  *tr = eps[0] + eps[4] + eps[8];
}

double orc_lin_energy(const double F[9], double lam, double mu) {
  double eps[9], tr, n2 = 0;
  lin_strain(F, eps, &tr);
  for (int k = 0; k < 9; ++k) n2 += eps[k] * eps[k];
  return mu * n2 + 0.5 * lam * tr * tr;
}

double orc_lin_stress(const double F[9], double lam, double mu, double P[9]) {
  double eps[9], tr;
  lin_strain(F, eps, &tr);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) P[3 * i + j] = 2 * mu * eps[3 * i + j] + (i == j ? lam * tr : 0.0);
  return orc_lin_energy(F, lam, mu);
}

void orc_lin_hessian(double lam, double mu, double H[81]) {
  for (int k = 0; k < 81; ++k) H[k] = 0.0;
  for (int a = 0; a < 9; ++a) H[a * 9 + a] = mu;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      H[(3 * i + j) * 9 + 3 * j + i] += mu;
      H[(3 * i + i) * 9 + 3 * j + j] += lam;
    }
}

/* 2D models, F row-major [F00, F01, F10, F11], Hessian 4x4 in that flattening.
 * Stable Neo-Hookean with Ndim = 2 (elast/stable_neohookean.hpp:23-107): δ the positive root of
 * (2t+1)δ² + (2t+2)δ − 2t = 0 with t = μ/λ, α = (1 − 1/(2+δ))μ/λ + 1,
 * Ψ = ½(μ(Ic − 2) + λ(J − α)² − μ ln(Ic + δ)) − ½(λ(1 − α)² − μ ln(2 + δ)),
 * P = λ(J − α)·cof F + μ(1 − 1/(Ic+δ))·F, cof F = [[F11, −F10], [−F01, F00]] (internal_physic.hpp:32-37). */
static double snh2_delta(double lam, double mu) {
  const double t = mu / lam;
  const double a = 2 * t + 1, b = 2 * t + 2, c = -2 * t;
  return (-b + sqrt(b * b - 4 * a * c)) / (2 * a);
}

double orc_snh2_energy(const double F[4], double lam, double mu) {
  const double del = snh2_delta(lam, mu);
  const double Ic = F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + F[3] * F[3];
  const double J = F[0] * F[3] - F[1] * F[2];
  const double alpha = (1 - 1 / (2 + del)) * mu / lam + 1;
  const double zero = 0.5 * (lam * (1 - alpha) * (1 - alpha) - mu * log(2 + del));
  return 0.5 * (mu * (Ic - 2) + lam * (J - alpha) * (J - alpha) - mu * log(Ic + del)) - zero;
}

double orc_snh2_stress(const double F[4], double lam, double mu, double P[4]) {
  const double del = snh2_delta(lam, mu);
  const double Ic = F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + F[3] * F[3];
  const double J = F[0] * F[3] - F[1] * F[2];
  const double alpha = (1 - 1 / (2 + del)) * mu / lam + 1;
  const double cof[4] = {F[3], -F[2], -F[1], F[0]};
  for (int i = 0; i < 4; ++i) P[i] = lam * (J - alpha) * cof[i] + mu * (1 - 1 / (Ic + del)) * F[i];
  return orc_snh2_energy(F, lam, mu);
}

/* H = μ(1 − 1/(Ic+δ))I + 2μ/(Ic+δ)² f fᵀ + λ g gᵀ + λ(J − α) ∂²J/∂F² (stable_neohookean.hpp:64-92, add_HJ
 * internal_physic.hpp:54-60) */
void orc_snh2_hessian(const double F[4], double lam, double mu, double H[16]) {
  const double del = snh2_delta(lam, mu);
  const double Ic = F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + F[3] * F[3];
  const double J = F[0] * F[3] - F[1] * F[2];
  const double alpha = (1 - 1 / (2 + del)) * mu / lam + 1;
  const double g[4] = {F[3], -F[2], -F[1], F[0]};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      H[4 * i + j] = (i == j ? mu * (1 - 1 / (Ic + del)) : 0.0) + 2 * mu / ((Ic + del) * (Ic + del)) * F[i] * F[j] +
                     lam * g[i] * g[j];
  const double s = lam * (J - alpha);
  H[3] += s;
  H[12] += s;
  H[6] -= s;
  H[9] -= s;
}

/* linear elasticity with Ndim = 2 (elast/linear.hpp:24-69) */
double orc_lin2_energy(const double F[4], double lam, double mu) {
  double n2 = 0;
  const double tr = F[0] + F[3] - 2;
  for (int a = 0; a < 2; ++a)
    for (int c = 0; c < 2; ++c) {
      const double e = 0.5 * (F[2 * a + c] + F[2 * c + a]) - (a == c ? 1.0 : 0.0);
      n2 += e * e;
    }
  return mu * n2 + 0.5 * lam * tr * tr;
}

double orc_lin2_stress(const double F[4], double lam, double mu, double P[4]) {
  const double tr = F[0] + F[3] - 2;
  for (int a = 0; a < 2; ++a)
    for (int c = 0; c < 2; ++c)
      P[2 * a + c] = 2 * mu * (0.5 * (F[2 * a + c] + F[2 * c + a]) - (a == c ? 1.0 : 0.0)) + (a == c ? lam * tr : 0.0);
  return orc_lin2_energy(F, lam, mu);
}

void orc_lin2_hessian(double lam, double mu, double H[16]) {
  for (int k = 0; k < 16; ++k) H[k] = 0.0;
  for (int a = 0; a < 4; ++a) H[5 * a] = mu;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      H[(2 * i + j) * 4 + 2 * j + i] += mu;
      H[(2 * i + i) * 4 + 2 * j + j] += lam;
    }
}

static void inv3(const double A[9], double Ai[9]) {
  const double d = det3(A);
  Ai[0] = (A[4] * A[8] - A[5] * A[7]) / d;
  Ai[1] = (A[2] * A[7] - A[1] * A[8]) / d;
  Ai[2] = (A[1] * A[5] - A[2] * A[4]) / d;
  Ai[3] = (A[5] * A[6] - A[3] * A[8]) / d;
  Ai[4] = (A[0] * A[8] - A[2] * A[6]) / d;
  Ai[5] = (A[2] * A[3] - A[0] * A[5]) / d;
  Ai[6] = (A[3] * A[7] - A[4] * A[6]) / d;
  Ai[7] = (A[1] * A[6] - A[0] * A[7]) / d;
  Ai[8] = (A[0] * A[4] - A[1] * A[3]) / d;
}

/* Dm = [x1−x0 | x2−x0 | x3−x0] (columns); B = Dm⁻¹ stored row-major: finite_elements/def_grad.hpp:229-247.
 * V = |det Dm| / 6: finite_elements/rest_vol.hpp:142-162. */
void orc_dminv(const double x0[3], const double x1[3], const double x2[3], const double x3[3], double B[9],
               double* V) {
  double Dm[9];
  for (int a = 0; a < 3; ++a) {
    Dm[a * 3 + 0] = x1[a] - x0[a];
    Dm[a * 3 + 1] = x2[a] - x0[a];
    Dm[a * 3 + 2] = x3[a] - x0[a];
  }
  inv3(Dm, B);
  if (V) *V = fabs(det3(Dm) / 6);
}

/* F = Ds·B with Ds columns x_k − x0: finite_elements/def_grad.hpp:270-288 */
void orc_def_grad(const double x0[3], const double x1[3], const double x2[3], const double x3[3],
                  const double B[9], double F[9]) {
  double Ds[9];
  for (int a = 0; a < 3; ++a) {
    Ds[a * 3 + 0] = x1[a] - x0[a];
    Ds[a * 3 + 1] = x2[a] - x0[a];
    Ds[a * 3 + 2] = x3[a] - x0[a];
  }
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) F[a * 3 + c] = Ds[a * 3 + 0] * B[0 * 3 + c] + Ds[a * 3 + 1] * B[1 * 3 + c] +
                                              Ds[a * 3 + 2] * B[2 * 3 + c];
}

/* ∂F/∂x, 9x12 row-major: pfpx[3a+c][3v+b] = δ_ab D[v][c], D[0][c] = −Σ_k B[k][c], D[v][c] = B[v−1][c].
 * elast/internal_physic.hpp:106-160 with the zero fill of finite_elements/def_grad.hpp:308-318 */
void orc_pfpx(const double B[9], double pfpx[108]) {
  memset(pfpx, 0, sizeof(double) * 108);
  double D[4][3];
  for (int c = 0; c < 3; ++c) {
    D[0][c] = -B[0 * 3 + c] - B[1 * 3 + c] - B[2 * 3 + c];
    for (int v = 1; v < 4; ++v) D[v][c] = B[(v - 1) * 3 + c];
  }
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c)
      for (int v = 0; v < 4; ++v) pfpx[(3 * a + c) * 12 + 3 * v + a] = D[v][c];
}

/* f_loc[v][b] = Σ_c P[b][c]·B[v−1][c] for v ≥ 1, f_loc[0] = −Σ_{v≥1} f_loc[v]:
 * finite_elements/time_step.hpp:73-89 (stress_element_to_vertex) */
void orc_local_force(const double P[9], const double B[9], double f[12]) {
  for (int j = 0; j < 3; ++j) {
    double tot = 0;
    for (int i = 1; i < 4; ++i) {
      const double val = P[j * 3 + 0] * B[(i - 1) * 3 + 0] + P[j * 3 + 1] * B[(i - 1) * 3 + 1] +
                         P[j * 3 + 2] * B[(i - 1) * 3 + 2];
      f[i * 3 + j] = val;
      tot += val;
    }
    f[0 * 3 + j] = -tot;
  }
}

/* K = pfpxᵀ · H · pfpx (12x12, not volume scaled): finite_elements/time_step.hpp:41-59 */
void orc_local_stiffness(const double H[81], const double B[9], double K[144]) {
  double pf[108], T[9 * 12];
  orc_pfpx(B, pf);
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 12; ++j) {
      double s = 0;
      for (int k = 0; k < 9; ++k) s += H[i * 9 + k] * pf[k * 12 + j];
      T[i * 12 + j] = s;
    }
  for (int i = 0; i < 12; ++i)
    for (int j = 0; j < 12; ++j) {
      double s = 0;
      for (int k = 0; k < 9; ++k) s += pf[k * 12 + i] * T[k * 12 + j];
      K[i * 12 + j] = s;
    }
}

/* ------------------------------------------------------------------------------------------------ */
/* problem level                                                                                    */

struct orc_ts {
  int nv, nt, n, nnz;
  double *X, *lam, *mu, *B, *V;
  int* cells;
  double rho, dt;
  double damping; /* velocity damping of step_next, in [0, 1) (0 after create, as the reference's default) */
  /* scalar CSR with the full 3x3 block pattern (finite_elements/global_composer.hpp:80-92) */
  int *rowptr, *colidx;
  int *brow, *bcol; /* block adjacency: brow[nv+1], bcol[nblocks] (sorted neighbour ids) */
  int* tet_kw;      /* per tet: 16 block positions kw (index of cell[j] within block row cell[i]) */
  int *inc_ptr, *inc;  /* vertex -> incident (tet*4 + corner), sorted */
  double* mass;     /* CSR values of M */
  unsigned char* dof;
  double *dbc, *u, *u_prev, *u_hat, *u_next, *vel, *accel, *g, *tmp;
  double *Kloc, *floc, *psi; /* per-tet scratch */
  double *Hvals;
  double rtol, tol_abs;
  int max_newton, pcg_maxit, ls_max;
  int polish; /* full Newton steps after the Newton loop (0: the reference's algorithm) */
  int model;  /* 0: stable Neo-Hookean, 1: linear elasticity */
  double* prec; /* NCG preconditioner factor L in the CSR order of the Hessian (NULL: none) */
  double* pd_vals; /* projective-dynamics matrix M + dt²κL (Hessian CSR order), built by orc_ts_build_pd */
  int lbfgs_pd;    /* L-BFGS middle step: 0 rest-Hessian diagonal, 1 solve with pd_vals, 2 solve with the Hessian */
  double pcg_rtol, ls_c, ls_beta;
  int lbfgs_m, lbfgs_maxit, lbfgs_ls_max;
  double lbfgs_c, lbfgs_beta;
  double* hrest_diag;
};

static int cmp_int(const void* a, const void* b) {
  const int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

/* deterministic blocked sum: fixed 4096-element chunks summed in parallel, chunk sums serially */
static double dot_n(int n, const double* a, const double* b) {
  const int CH = 4096;
  const int nch = (n + CH - 1) / CH;
  double* part = (double*)malloc(sizeof(double) * (nch > 0 ? nch : 1));
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nch; ++c) {
    double s = 0;
    const int lo = c * CH, hi = lo + CH < n ? lo + CH : n;
    for (int i = lo; i < hi; ++i) s += a[i] * b[i];
    part[c] = s;
  }
  double s = 0;
  for (int c = 0; c < nch; ++c) s += part[c];
  free(part);
  return s;
}

orc_ts* orc_ts_create(const double* X, int nv, const int* cells, int nt, double dt, const double* lam,
                      const double* mu, double rho) {
  orc_ts* ts = (orc_ts*)calloc(1, sizeof(orc_ts));
  ts->nv = nv;
  ts->nt = nt;
  ts->n = 3 * nv;
  ts->dt = dt;
  ts->rho = rho;
  ts->X = (double*)malloc(sizeof(double) * 3 * nv);
  memcpy(ts->X, X, sizeof(double) * 3 * nv);
  ts->cells = (int*)malloc(sizeof(int) * 4 * nt);
  memcpy(ts->cells, cells, sizeof(int) * 4 * nt);
  ts->lam = (double*)malloc(sizeof(double) * nt);
  ts->mu = (double*)malloc(sizeof(double) * nt);
  memcpy(ts->lam, lam, sizeof(double) * nt);
  memcpy(ts->mu, mu, sizeof(double) * nt);
  ts->B = (double*)malloc(sizeof(double) * 9 * nt);
  ts->V = (double*)malloc(sizeof(double) * nt);
  /* reset(): dminv_task, rest_vol_task (finite_elements/time_step.hpp:426-431) */
#pragma omp parallel for schedule(static)
  for (int e = 0; e < nt; ++e) {
    const int* c = cells + 4 * e;
    orc_dminv(X + 3 * c[0], X + 3 * c[1], X + 3 * c[2], X + 3 * c[3], ts->B + 9 * e, ts->V + e);
  }
  /* block adjacency (the CSR topology of finite_elements/global_composer.hpp:80-92, in 3x3 blocks) */
  int* deg = (int*)calloc(nv + 1, sizeof(int));
  for (int e = 0; e < nt; ++e)
    for (int i = 0; i < 4; ++i) deg[cells[4 * e + i]] += 4;
  int* cand_ptr = (int*)malloc(sizeof(int) * (nv + 1));
  cand_ptr[0] = 0;
  for (int v = 0; v < nv; ++v) cand_ptr[v + 1] = cand_ptr[v] + deg[v];
  int* cand = (int*)malloc(sizeof(int) * (cand_ptr[nv] > 0 ? cand_ptr[nv] : 1));
  int* fill = (int*)calloc(nv, sizeof(int));
  for (int e = 0; e < nt; ++e)
    for (int i = 0; i < 4; ++i) {
      const int v = cells[4 * e + i];
      for (int j = 0; j < 4; ++j) cand[cand_ptr[v] + fill[v]++] = cells[4 * e + j];
    }
  ts->brow = (int*)malloc(sizeof(int) * (nv + 1));
  ts->brow[0] = 0;
  for (int v = 0; v < nv; ++v) {
    int* lst = cand + cand_ptr[v];
    const int m = cand_ptr[v + 1] - cand_ptr[v];
    qsort(lst, m, sizeof(int), cmp_int);
    int u = 0;
    for (int k = 0; k < m; ++k)
      if (k == 0 || lst[k] != lst[k - 1]) lst[u++] = lst[k];
    fill[v] = u;
    ts->brow[v + 1] = ts->brow[v] + u;
  }
  ts->bcol = (int*)malloc(sizeof(int) * (ts->brow[nv] > 0 ? ts->brow[nv] : 1));
  for (int v = 0; v < nv; ++v) memcpy(ts->bcol + ts->brow[v], cand + cand_ptr[v], sizeof(int) * fill[v]);
  free(cand);
  free(cand_ptr);
  free(fill);
  free(deg);
  /* scalar CSR: row 3v+a holds columns 3w+b for every block neighbour w, ascending */
  ts->rowptr = (int*)malloc(sizeof(int) * (3 * nv + 1));
  ts->rowptr[0] = 0;
  for (int v = 0; v < nv; ++v) {
    const int len = 3 * (ts->brow[v + 1] - ts->brow[v]);
    for (int a = 0; a < 3; ++a) ts->rowptr[3 * v + a + 1] = ts->rowptr[3 * v + a] + len;
  }
  ts->nnz = ts->rowptr[3 * nv];
  ts->colidx = (int*)malloc(sizeof(int) * (ts->nnz > 0 ? ts->nnz : 1));
  for (int v = 0; v < nv; ++v)
    for (int a = 0; a < 3; ++a) {
      int p = ts->rowptr[3 * v + a];
      for (int k = ts->brow[v]; k < ts->brow[v + 1]; ++k)
        for (int b = 0; b < 3; ++b) ts->colidx[p++] = 3 * ts->bcol[k] + b;
    }
  /* per-tet block positions (the hessian gather map, finite_elements/global_composer.hpp:102-119) */
  ts->tet_kw = (int*)malloc(sizeof(int) * 16 * nt);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < nt; ++e)
    for (int i = 0; i < 4; ++i) {
      const int v = cells[4 * e + i];
      for (int j = 0; j < 4; ++j) {
        const int w = cells[4 * e + j];
        int lo = ts->brow[v], hi = ts->brow[v + 1] - 1;
        while (lo < hi) {
          const int mid = (lo + hi) / 2;
          if (ts->bcol[mid] < w) lo = mid + 1; else hi = mid;
        }
        ts->tet_kw[16 * e + 4 * i + j] = lo - ts->brow[v];
      }
    }
  /* vertex -> incident (tet, corner), in tet order (the force gather map, global_composer.hpp:36-66) */
  ts->inc_ptr = (int*)calloc(nv + 1, sizeof(int));
  for (int e = 0; e < nt; ++e)
    for (int i = 0; i < 4; ++i) ts->inc_ptr[cells[4 * e + i] + 1]++;
  for (int v = 0; v < nv; ++v) ts->inc_ptr[v + 1] += ts->inc_ptr[v];
  ts->inc = (int*)malloc(sizeof(int) * (4 * nt > 0 ? 4 * nt : 1));
  int* pos = (int*)malloc(sizeof(int) * (nv > 0 ? nv : 1));
  for (int v = 0; v < nv; ++v) pos[v] = ts->inc_ptr[v];
  for (int e = 0; e < nt; ++e)
    for (int i = 0; i < 4; ++i) ts->inc[pos[cells[4 * e + i]]++] = 4 * e + i;
  free(pos);
  /* consistent mass: M_e[i][j] = ρ |det Dm| (i==j ? 1/60 : 1/120) (finite_elements/bilinears.hpp:140-148,
   * 222-251), M[3I+a][3J+b] = δ_ab Σ_e M_e[i][j] with explicit zeros (finite_elements/time_step.hpp:437-458) */
  ts->mass = (double*)calloc(ts->nnz > 0 ? ts->nnz : 1, sizeof(double));
#pragma omp parallel for schedule(dynamic, 64)
  for (int v = 0; v < nv; ++v) {
    for (int q = ts->inc_ptr[v]; q < ts->inc_ptr[v + 1]; ++q) {
      const int e = ts->inc[q] / 4, i = ts->inc[q] % 4;
      const double detj = 6 * ts->V[e];
      for (int j = 0; j < 4; ++j) {
        const double m = rho * detj * (i == j ? 1.0 / 60 : 1.0 / 120);
        const int kw = ts->tet_kw[16 * e + 4 * i + j];
        for (int a = 0; a < 3; ++a) ts->mass[ts->rowptr[3 * v + a] + 3 * kw + a] += m;
      }
    }
  }
  const int n = ts->n;
  ts->dof = (unsigned char*)calloc(n, 1);
  double** vecs[] = {&ts->dbc, &ts->u, &ts->u_prev, &ts->u_hat, &ts->u_next, &ts->vel, &ts->accel, &ts->g,
                     &ts->tmp};
  for (unsigned k = 0; k < sizeof(vecs) / sizeof(vecs[0]); ++k) *vecs[k] = (double*)calloc(n, sizeof(double));
  ts->Kloc = NULL;
  ts->floc = (double*)malloc(sizeof(double) * 12 * (nt > 0 ? nt : 1));
  ts->psi = (double*)malloc(sizeof(double) * (nt > 0 ? nt : 1));
  ts->Hvals = (double*)calloc(ts->nnz > 0 ? ts->nnz : 1, sizeof(double));
  ts->rtol = 1e-2; /* finite_elements/time_step.hpp:706 */
  ts->tol_abs = ts->rtol * orc_ts_total_mass(ts);
  ts->max_newton = 10;
  ts->pcg_rtol = 1e-8;
  ts->pcg_maxit = 10000;
  ts->ls_c = 0.3; /* stepper/backward_euler.hpp:22 */
  ts->ls_beta = 0.5;
  ts->ls_max = 20;
  ts->polish = 0;
  ts->model = 0;
  ts->prec = NULL;
  ts->pd_vals = NULL;
  ts->lbfgs_pd = 0;
  ts->lbfgs_m = 8;
  ts->lbfgs_maxit = 1000;
  ts->lbfgs_c = 1e-4;
  ts->lbfgs_beta = 0.5;
  ts->lbfgs_ls_max = 30;
  return ts;
}

void orc_ts_destroy(orc_ts* ts) {
  if (!ts) return;
  void* ptrs[] = {ts->X, ts->lam, ts->mu, ts->B, ts->V, ts->cells, ts->rowptr, ts->colidx, ts->brow, ts->bcol,
                  ts->tet_kw, ts->inc_ptr, ts->inc, ts->mass, ts->dof, ts->dbc, ts->u, ts->u_prev, ts->u_hat,
                  ts->u_next, ts->vel, ts->accel, ts->g, ts->tmp, ts->Kloc, ts->floc, ts->psi, ts->Hvals,
                  ts->hrest_diag, ts->prec, ts->pd_vals};
  for (unsigned k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); ++k) free(ptrs[k]);
  free(ts);
}

int orc_ts_nnz(const orc_ts* ts) { return ts->nnz; }
void orc_ts_csr_pattern(const orc_ts* ts, int* rowptr, int* colidx) {
  memcpy(rowptr, ts->rowptr, sizeof(int) * (ts->n + 1));
  memcpy(colidx, ts->colidx, sizeof(int) * ts->nnz);
}
void orc_ts_mass_values(const orc_ts* ts, double* vals) { memcpy(vals, ts->mass, sizeof(double) * ts->nnz); }

/* asum(M.values)/3 = total mass (finite_elements/time_step.hpp:490) */
double orc_ts_total_mass(const orc_ts* ts) {
  double s = 0;
  for (int k = 0; k < ts->nnz; ++k) s += fabs(ts->mass[k]);
  return s / 3;
}
void orc_ts_rest_data(const orc_ts* ts, double* B, double* V) {
  if (B) memcpy(B, ts->B, sizeof(double) * 9 * ts->nt);
  if (V) memcpy(V, ts->V, sizeof(double) * ts->nt);
}

void orc_ts_set_dirichlet(orc_ts* ts, const unsigned char* dof_type, const double* dbc_values) {
  memcpy(ts->dof, dof_type, ts->n);
  memcpy(ts->dbc, dbc_values, sizeof(double) * ts->n);
}
void orc_ts_set_ext_accel(orc_ts* ts, const double* accel) { memcpy(ts->accel, accel, sizeof(double) * ts->n); }
void orc_ts_set_threshold(orc_ts* ts, double rtol) {
  ts->rtol = rtol;
  ts->tol_abs = rtol * orc_ts_total_mass(ts);
}
void orc_ts_set_state(orc_ts* ts, const double* u, const double* u_prev, const double* vel) {
  if (u) memcpy(ts->u, u, sizeof(double) * ts->n);
  if (u_prev) memcpy(ts->u_prev, u_prev, sizeof(double) * ts->n);
  if (vel) memcpy(ts->vel, vel, sizeof(double) * ts->n);
}
void orc_ts_get_state(const orc_ts* ts, double* u, double* u_prev, double* vel) {
  if (u) memcpy(u, ts->u, sizeof(double) * ts->n);
  if (u_prev) memcpy(u_prev, ts->u_prev, sizeof(double) * ts->n);
  if (vel) memcpy(vel, ts->vel, sizeof(double) * ts->n);
}
void orc_ts_set_vectors(orc_ts* ts, const double* u_next, const double* u_hat) {
  if (u_next) memcpy(ts->u_next, u_next, sizeof(double) * ts->n);
  if (u_hat) memcpy(ts->u_hat, u_hat, sizeof(double) * ts->n);
}
void orc_ts_get_vectors(const orc_ts* ts, double* u_next, double* u_hat, double* forces) {
  if (u_next) memcpy(u_next, ts->u_next, sizeof(double) * ts->n);
  if (u_hat) memcpy(u_hat, ts->u_hat, sizeof(double) * ts->n);
  if (forces) memcpy(forces, ts->g, sizeof(double) * ts->n);
}
void orc_ts_set_solver(orc_ts* ts, int max_newton, double pcg_rtol, int pcg_maxit, double ls_c, double ls_beta,
                       int ls_max) {
  ts->max_newton = max_newton;
  ts->pcg_rtol = pcg_rtol;
  ts->pcg_maxit = pcg_maxit;
  ts->ls_c = ls_c;
  ts->ls_beta = ls_beta;
  ts->ls_max = ls_max;
}
void orc_ts_set_polish(orc_ts* ts, int polish) { ts->polish = polish; }
/* tet_ncg_cuda_ext_ai::set_matrix (bindings/src/fem_cuda.cu:204): L in the Hessian's CSR order; NULL removes it */
void orc_ts_set_precond(orc_ts* ts, const double* values) {
  free(ts->prec);
  ts->prec = NULL;
  if (!values) return;
  ts->prec = (double*)malloc(sizeof(double) * (ts->nnz > 0 ? ts->nnz : 1));
  memcpy(ts->prec, values, sizeof(double) * ts->nnz);
}

void orc_ts_set_model(orc_ts* ts, int model) {
  ts->model = model;
  if (ts->hrest_diag) {
    free(ts->hrest_diag);
    ts->hrest_diag = NULL;
  }
}

void orc_ts_set_lbfgs(orc_ts* ts, int m, int max_iter, double ls_c, double ls_beta, int ls_max) {
  ts->lbfgs_m = m;
  ts->lbfgs_maxit = max_iter;
  ts->lbfgs_c = ls_c;
  ts->lbfgs_beta = ls_beta;
  ts->lbfgs_ls_max = ls_max;
}

static void tet_positions(const orc_ts* ts, const double* u, int e, double x[4][3]) {
  for (int i = 0; i < 4; ++i) {
    const int v = ts->cells[4 * e + i];
    for (int a = 0; a < 3; ++a) x[i][a] = ts->X[3 * v + a] + u[3 * v + a];
  }
}

void orc_ts_elements(const orc_ts* ts, const double* u, double* psi, double* P, double* H, double* Hp,
                     double* K, int make_spsd) {
#pragma omp parallel for schedule(static)
  for (int e = 0; e < ts->nt; ++e) {
    double x[4][3], F[9], Pl[9], Hl[81], Hpl[81];
    tet_positions(ts, u, e, x);
    orc_def_grad(x[0], x[1], x[2], x[3], ts->B + 9 * e, F);
    const double ps = ts->model ? orc_lin_stress(F, ts->lam[e], ts->mu[e], Pl) : orc_snh_stress(F, ts->lam[e], ts->mu[e], Pl);
    if (psi) psi[e] = ps;
    if (P) memcpy(P + 9 * e, Pl, sizeof(Pl));
    if (H || Hp || K) {
      if (ts->model)
        orc_lin_hessian(ts->lam[e], ts->mu[e], Hl);
      else
        orc_snh_hessian(F, ts->lam[e], ts->mu[e], Hl);
      if (H) memcpy(H + 81 * e, Hl, sizeof(Hl));
      if (Hp || (K && make_spsd)) {
        orc_spd_project9(Hl, 1e-3, Hpl);
        if (Hp) memcpy(Hp + 81 * e, Hpl, sizeof(Hpl));
      }
      if (K) orc_local_stiffness(make_spsd ? Hpl : Hl, ts->B + 9 * e, K + 144 * e);
    }
  }
}

static void csr_spmv(int nrows, const int* rowptr, const int* colidx, const double* vals, const double* x,
                     double* y) {
#pragma omp parallel for schedule(static)
  for (int r = 0; r < nrows; ++r) {
    double s = 0;
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) s += vals[k] * x[colidx[k]];
    y[r] = s;
  }
}

void orc_ts_spmv(const orc_ts* ts, const double* vals, const double* x, double* y) {
  csr_spmv(ts->n, ts->rowptr, ts->colidx, vals, x, y);
}

/* update_energy_and_gradients: finite_elements/time_step.hpp:577-624.  The unfiltered mass is used for the
 * inertia term (mass_bl_, :591); forces gathered with weight V and alpha dt² (:606-613); DBC grads zeroed
 * (:615-617); E = ½ tempᵀ M temp + dt² Σ V Ψ (:592, :621). */
double orc_ts_energy_grad(orc_ts* ts, const double* u, double* g) {
  const int n = ts->n, nt = ts->nt, nv = ts->nv;
  const double dt2 = ts->dt * ts->dt;
  double* temp = ts->tmp;
  for (int i = 0; i < n; ++i) temp[i] = u[i] - ts->u_hat[i];
  csr_spmv(n, ts->rowptr, ts->colidx, ts->mass, temp, g);
  const double inertia = 0.5 * dot_n(n, temp, g);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < nt; ++e) {
    double x[4][3], F[9], P[9];
    tet_positions(ts, u, e, x);
    orc_def_grad(x[0], x[1], x[2], x[3], ts->B + 9 * e, F);
    ts->psi[e] = ts->model ? orc_lin_stress(F, ts->lam[e], ts->mu[e], P) : orc_snh_stress(F, ts->lam[e], ts->mu[e], P);
    orc_local_force(P, ts->B + 9 * e, ts->floc + 12 * e);
  }
#pragma omp parallel for schedule(dynamic, 256)
  for (int v = 0; v < nv; ++v) {
    double acc[3] = {0, 0, 0};
    for (int q = ts->inc_ptr[v]; q < ts->inc_ptr[v + 1]; ++q) {
      const int e = ts->inc[q] / 4, i = ts->inc[q] % 4;
      for (int a = 0; a < 3; ++a) acc[a] += ts->V[e] * ts->floc[12 * e + 3 * i + a];
    }
    for (int a = 0; a < 3; ++a) g[3 * v + a] += dt2 * acc[a];
  }
  for (int i = 0; i < n; ++i)
    if (ts->dof[i]) g[i] = 0; /* zero_grads_impl, finite_elements/boundary.hpp:45-51 */
  const double elast = dot_n(nt, ts->psi, ts->V) * dt2;
  return inertia + elast;
}

/* update_hessian: finite_elements/time_step.hpp:626-658.  H = M + dt² Σ V K; DBC rows and columns set to
 * the identity (identity_hessian_impl, finite_elements/boundary.hpp:53-67). */
void orc_ts_hessian(orc_ts* ts, const double* u, int make_spsd, int filter_dbc, double* vals) {
  const int nt = ts->nt, nv = ts->nv;
  const double dt2 = ts->dt * ts->dt;
  if (!ts->Kloc) ts->Kloc = (double*)malloc(sizeof(double) * 144 * (nt > 0 ? nt : 1));
  orc_ts_elements(ts, u, NULL, NULL, NULL, NULL, ts->Kloc, make_spsd);
  memcpy(vals, ts->mass, sizeof(double) * ts->nnz);
#pragma omp parallel for schedule(dynamic, 64)
  for (int v = 0; v < nv; ++v) {
    for (int q = ts->inc_ptr[v]; q < ts->inc_ptr[v + 1]; ++q) {
      const int e = ts->inc[q] / 4, i = ts->inc[q] % 4;
      const double* Ke = ts->Kloc + 144 * e;
      for (int j = 0; j < 4; ++j) {
        const int kw = ts->tet_kw[16 * e + 4 * i + j];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            vals[ts->rowptr[3 * v + a] + 3 * kw + b] += dt2 * ts->V[e] * Ke[(3 * i + a) * 12 + 3 * j + b];
      }
    }
  }
  if (filter_dbc) {
    for (int r = 0; r < ts->n; ++r)
      for (int k = ts->rowptr[r]; k < ts->rowptr[r + 1]; ++k) {
        const int c = ts->colidx[k];
        if (ts->dof[r] || ts->dof[c]) vals[k] = (r == c) ? 1.0 : 0.0;
      }
  }
}

/* Jacobi-preconditioned CG (Hestenes–Stiefel), x0 = 0, stop when ‖r‖₂ ≤ rtol·‖b‖₂ (checked before each
 * iteration).  Stands in for mathprim cg + diagonal_preconditioner (tests/fem_cu/main.cu:52-54). */
int orc_pcg(int n, const int* rowptr, const int* colidx, const double* vals, const double* b, double* x,
            double rtol, int maxit, double* out_rel_resid) {
  double* dinv = (double*)malloc(sizeof(double) * n);
  double* r = (double*)malloc(sizeof(double) * n);
  double* z = (double*)malloc(sizeof(double) * n);
  double* p = (double*)malloc(sizeof(double) * n);
  double* q = (double*)malloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) {
    double d = 0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
      if (colidx[k] == i) d = vals[k];
    dinv[i] = 1.0 / d;
    x[i] = 0;
    r[i] = b[i];
    z[i] = dinv[i] * r[i];
    p[i] = z[i];
  }
  double rz = dot_n(n, r, z);
  const double bb = dot_n(n, b, b);
  double rr = bb;
  int it = 0;
  for (; it < maxit; ++it) {
    if (rr <= rtol * rtol * bb) break;
    csr_spmv(n, rowptr, colidx, vals, p, q);
    const double pq = dot_n(n, p, q);
    const double alpha = rz / pq;
    for (int i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      z[i] = dinv[i] * r[i];
    }
    const double rz_new = dot_n(n, r, z);
    rr = dot_n(n, r, r);
    const double beta = rz_new / rz;
    rz = rz_new;
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
  if (out_rel_resid) *out_rel_resid = bb > 0 ? sqrt(rr / bb) : 0.0;
  free(dinv);
  free(r);
  free(z);
  free(p);
  free(q);
  return it;
}

/* prepare_step: û = u + dt v + dt² a; û[DBC] = dbc; u_next = û (finite_elements/time_step.hpp:499-518) */
void orc_ts_prepare_step(orc_ts* ts) {
  const double dt = ts->dt;
  for (int i = 0; i < ts->n; ++i) {
    double uh = ts->u[i];
    uh += dt * ts->vel[i];
    uh += dt * dt * ts->accel[i];
    if (ts->dof[i]) uh = ts->dbc[i];
    ts->u_hat[i] = uh;
    ts->u_next[i] = uh;
  }
}

/* Newton with backtracking Armijo line search (stepper/backward_euler.hpp:15-33: tol_grad = tol_abs,
 * armijo threshold 0.3; at most max_newton iterations per the benchmark configuration), followed by the
 * final update_energy_and_gradients of compute_step (finite_elements/time_step.hpp:520-528). */
double orc_ts_compute_step(orc_ts* ts, double* stats) {
  const int n = ts->n;
  double* u = ts->u_next;
  double* d = (double*)malloc(sizeof(double) * n);
  double* mg = (double*)malloc(sizeof(double) * n);
  double* ut = (double*)malloc(sizeof(double) * n);
  double* gt = (double*)malloc(sizeof(double) * n);
  int newton = 0, pcg_total = 0, evals = 0, converged = 0, ls_fail = 0, ls_rejected = 0;
  double last_change = 0.0;
  double E = orc_ts_energy_grad(ts, u, ts->g);
  evals++;
  for (; newton < ts->max_newton;) {
    const double gn = sqrt(dot_n(n, ts->g, ts->g));
    if (gn < ts->tol_abs) {
      converged = 1;
      break;
    }
    orc_ts_hessian(ts, u, 1, 1, ts->Hvals);
    for (int i = 0; i < n; ++i) mg[i] = -ts->g[i];
    pcg_total += orc_pcg(n, ts->rowptr, ts->colidx, ts->Hvals, mg, d, ts->pcg_rtol, ts->pcg_maxit, NULL);
    const double gd = dot_n(n, ts->g, d);
    double alpha = 1.0, Et = 0;
    int ok = 0;
    for (int ls = 0; ls < ts->ls_max; ++ls) {
      for (int i = 0; i < n; ++i) ut[i] = u[i] + alpha * d[i];
      Et = orc_ts_energy_grad(ts, ut, gt);
      evals++;
      if (Et <= E + ts->ls_c * alpha * gd) {
        ok = 1;
        break;
      }
      alpha *= ts->ls_beta;
    }
    newton++;
    if (!ok) {
      ls_fail++;
      /* every trial failed and the smallest still raises the energy: keep the iterate and stop (the same rule
       * as the GPU solver; the behaviour of mathprim's backtracking_linesearcher on failure is not known) */
      if (!(Et <= E)) {
        ls_rejected = 1;
        break;
      }
    }
    memcpy(u, ut, sizeof(double) * n);
    memcpy(ts->g, gt, sizeof(double) * n);
    last_change = fabs(E - Et);
    E = Et;
  }
  /* optional polish: full Newton steps (α = 1) with the unprojected Hessian, positive definite at the minimum
   * the loop has converged to, so they converge quadratically and take the solution to roundoff accuracy where
   * the Armijo test can no longer resolve energy decreases (extension for finite-difference checks; not in the
   * reference) */
  for (int k = 0; k < ts->polish; ++k) {
    orc_ts_hessian(ts, u, 0, 1, ts->Hvals);
    for (int i = 0; i < n; ++i) mg[i] = -ts->g[i];
    pcg_total += orc_pcg(n, ts->rowptr, ts->colidx, ts->Hvals, mg, d, ts->pcg_rtol, ts->pcg_maxit, NULL);
    for (int i = 0; i < n; ++i) u[i] += d[i];
    orc_ts_energy_grad(ts, u, ts->g);
    evals++;
  }
  /* compute_step: final update_energy_and_gradients + ‖forces‖ */
  orc_ts_energy_grad(ts, u, ts->g);
  const double gn = sqrt(dot_n(n, ts->g, ts->g));
  if (stats) {
    stats[0] = newton;
    stats[1] = pcg_total;
    stats[2] = evals;
    stats[3] = converged;
    stats[4] = gn;
    stats[5] = ls_fail;
    stats[6] = last_change;
    stats[7] = ls_rejected;
  }
  free(d);
  free(mg);
  free(ut);
  free(gt);
  return gn;
}

/* diagonal of the SPD-projected, DBC-filtered Hessian at rest (u = 0): the L-BFGS initial scaling H0 */
void orc_ts_rest_hessian_diag(orc_ts* ts, double* diag) {
  double* zero = (double*)calloc(ts->n, sizeof(double));
  double* vals = (double*)malloc(sizeof(double) * ts->nnz);
  orc_ts_hessian(ts, zero, 1, 1, vals);
  for (int r = 0; r < ts->n; ++r)
    for (int k = ts->rowptr[r]; k < ts->rowptr[r + 1]; ++k)
      if (ts->colidx[k] == r) diag[r] = vals[k];
  free(zero);
  free(vals);
}

/* L-BFGS (Nocedal & Wright, Numerical Optimization 2nd ed., Alg. 7.4 two-loop recursion, Alg. 7.5 driver)
 * with H0 = diag(H_rest)⁻¹ plugged in where the reference's l_bfgs_preconditioner::apply_impl(z, q, s, y)
 * hook sits (stepper/projective_dynamics.hpp:15-35), backtracking Armijo line search, stopping when
 * ‖g‖₂ < tol_abs (stepper/backward_euler.hpp:48-60). */
/* Projective-dynamics matrix of time_step_solver_pd::reset_impl (stepper/projective_dynamics.hpp:51-87):
 * A = M + dt²·κ·K on every dof, M the consistent mass ρV(1+δ_ij)/20 (mass_integrator, bilinears.hpp:140-148)
 * and K the P1 Laplacian V·∇N_i·∇N_j (laplace_integrator, bilinears.hpp:158-167, 260-264) with the diffusivity κ
 * of the uniform material (λ + 2μ, elast/basic_elast.hpp:225), Dirichlet rows and columns set to the identity. */
void orc_ts_build_pd(orc_ts* ts, double kappa) {
  const int nv = ts->nv;
  const double c = ts->dt * ts->dt * kappa;
  if (!ts->pd_vals) ts->pd_vals = (double*)malloc(sizeof(double) * (ts->nnz > 0 ? ts->nnz : 1));
  memcpy(ts->pd_vals, ts->mass, sizeof(double) * ts->nnz);
  for (int v = 0; v < nv; ++v)
    for (int q = ts->inc_ptr[v]; q < ts->inc_ptr[v + 1]; ++q) {
      const int e = ts->inc[q] / 4, i = ts->inc[q] % 4;
      const double* B = ts->B + 9 * e;
      double D[4][3];
      for (int k = 0; k < 3; ++k) {
        D[0][k] = -B[0 * 3 + k] - B[1 * 3 + k] - B[2 * 3 + k];
        for (int r = 1; r < 4; ++r) D[r][k] = B[(r - 1) * 3 + k];
      }
      for (int j = 0; j < 4; ++j) {
        const double lap = ts->V[e] * (D[i][0] * D[j][0] + D[i][1] * D[j][1] + D[i][2] * D[j][2]);
        const int kw = ts->tet_kw[16 * e + 4 * i + j];
        for (int a = 0; a < 3; ++a) ts->pd_vals[ts->rowptr[3 * v + a] + 3 * kw + a] += c * lap;
      }
    }
  for (int r = 0; r < ts->n; ++r)
    for (int k = ts->rowptr[r]; k < ts->rowptr[r + 1]; ++k) {
      const int col = ts->colidx[k];
      if (ts->dof[r] || ts->dof[col]) ts->pd_vals[k] = (r == col) ? 1.0 : 0.0;
    }
}
void orc_ts_pd_values(const orc_ts* ts, double* vals) {
  if (ts->pd_vals) memcpy(vals, ts->pd_vals, sizeof(double) * ts->nnz);
}

/* time_step_solver_pd (stepper/projective_dynamics.hpp:36-50): L-BFGS whose two-loop middle step solves with the
 * projective-dynamics matrix (a direct solver in the reference; here Jacobi-PCG to a relative residual of 1e-12). */
/* time_step_solver_h0 (stepper/projective_dynamics.hpp:91-148): L-BFGS whose two-loop middle step solves with the
 * SPD-projected, Dirichlet-filtered Hessian at the start of the step (factorised in the reference). */
double orc_ts_compute_step_lbfgs_h0(orc_ts* ts, double* stats) {
  ts->lbfgs_pd = 2;
  const double r = orc_ts_compute_step_lbfgs(ts, stats);
  ts->lbfgs_pd = 0;
  return r;
}

double orc_ts_compute_step_lbfgs_pd(orc_ts* ts, double* stats) {
  if (!ts->pd_vals) return -1.0;
  ts->lbfgs_pd = 1;
  const double r = orc_ts_compute_step_lbfgs(ts, stats);
  ts->lbfgs_pd = 0;
  return r;
}

double orc_ts_compute_step_lbfgs(orc_ts* ts, double* stats) {
  const int n = ts->n, m = ts->lbfgs_m;
  if (!ts->hrest_diag) {
    ts->hrest_diag = (double*)malloc(sizeof(double) * n);
    orc_ts_rest_hessian_diag(ts, ts->hrest_diag);
  }
  double* u = ts->u_next;
  double* S = (double*)calloc((size_t)m * n, sizeof(double));
  double* Y = (double*)calloc((size_t)m * n, sizeof(double));
  double* rho = (double*)calloc(m, sizeof(double));
  double* a = (double*)calloc(m, sizeof(double));
  double* q = (double*)malloc(sizeof(double) * n);
  double* ut = (double*)malloc(sizeof(double) * n);
  double* gt = (double*)malloc(sizeof(double) * n);
  int it = 0, evals = 0, converged = 0, ls_fail = 0, hist = 0, head = 0, pd_total = 0;
  if (ts->lbfgs_pd == 2) orc_ts_hessian(ts, u, 1, 1, ts->Hvals);
  double E = orc_ts_energy_grad(ts, u, ts->g);
  evals++;
  for (; it < ts->lbfgs_maxit;) {
    const double gn = sqrt(dot_n(n, ts->g, ts->g));
    if (gn < ts->tol_abs) {
      converged = 1;
      break;
    }
    memcpy(q, ts->g, sizeof(double) * n);
    for (int k = 0; k < hist; ++k) {
      const int idx = (head - 1 - k + m) % m;
      a[idx] = rho[idx] * dot_n(n, S + (size_t)idx * n, q);
      const double* y = Y + (size_t)idx * n;
      for (int i = 0; i < n; ++i) q[i] -= a[idx] * y[i];
    }
    if (ts->lbfgs_pd) { /* r = A⁻¹ q, A the PD matrix (1) or the Hessian at the start of the step (2) */
      pd_total += orc_pcg(n, ts->rowptr, ts->colidx, ts->lbfgs_pd == 2 ? ts->Hvals : ts->pd_vals, q, ut, 1e-12,
                          100000, NULL);
      memcpy(q, ut, sizeof(double) * n);
    } else {
      for (int i = 0; i < n; ++i) q[i] /= ts->hrest_diag[i];
    }
    for (int k = hist - 1; k >= 0; --k) {
      const int idx = (head - 1 - k + m) % m;
      const double bb = rho[idx] * dot_n(n, Y + (size_t)idx * n, q);
      const double* s = S + (size_t)idx * n;
      for (int i = 0; i < n; ++i) q[i] += s[i] * (a[idx] - bb);
    }
    for (int i = 0; i < n; ++i) q[i] = -q[i]; /* direction d */
    const double gd = dot_n(n, ts->g, q);
    double alpha = 1.0, Et = 0;
    int ok = 0;
    for (int ls = 0; ls < ts->lbfgs_ls_max; ++ls) {
      for (int i = 0; i < n; ++i) ut[i] = u[i] + alpha * q[i];
      Et = orc_ts_energy_grad(ts, ut, gt);
      evals++;
      if (Et <= E + ts->lbfgs_c * alpha * gd) {
        ok = 1;
        break;
      }
      alpha *= ts->lbfgs_beta;
    }
    if (!ok) ls_fail++;
    double* s = S + (size_t)head * n;
    double* y = Y + (size_t)head * n;
    for (int i = 0; i < n; ++i) {
      s[i] = ut[i] - u[i];
      y[i] = gt[i] - ts->g[i];
    }
    const double ys = dot_n(n, y, s);
    if (ys > 1e-300) {
      rho[head] = 1.0 / ys;
      head = (head + 1) % m;
      if (hist < m) hist++;
    }
    memcpy(u, ut, sizeof(double) * n);
    memcpy(ts->g, gt, sizeof(double) * n);
    E = Et;
    it++;
  }
  orc_ts_energy_grad(ts, u, ts->g);
  const double gn = sqrt(dot_n(n, ts->g, ts->g));
  if (stats) {
    stats[0] = it;
    stats[1] = pd_total;
    stats[2] = evals;
    stats[3] = converged;
    stats[4] = gn;
    stats[5] = ls_fail;
  }
  free(S);
  free(Y);
  free(rho);
  free(a);
  free(q);
  free(ut);
  free(gt);
  return gn;
}

/* z = L·(Lᵀ·g) with L in CSR (the preconditioner of time_step_solver_ncg_with_ext_prec,
 * stepper/nonlinear_cg.hpp:54-59: gemv with the transpose, then gemv): a = Lᵀg, out = L a */
static void prec_apply(const orc_ts* ts, const double* g, double* a, double* out) {
  const int n = ts->n;
  for (int i = 0; i < n; ++i) a[i] = 0.0;
  for (int r = 0; r < n; ++r)
    for (int k = ts->rowptr[r]; k < ts->rowptr[r + 1]; ++k) a[ts->colidx[k]] += ts->prec[k] * g[r];
  csr_spmv(n, ts->rowptr, ts->colidx, ts->prec, a, out);
}

/* Nonlinear CG stepper (stepper/nonlinear_cg.hpp:9-37: mathprim ncg_optimizer with the Dai-Kou strategy,
 * learning rate 1e-2, backtracking line search, at most 1000 iterations, tol = threshold·mass).  mathprim is not
 * available, so this follows the published method: Dai & Kou, "A nonlinear conjugate gradient algorithm with an
 * optimal property and an improved Wolfe line search", SIAM J. Optim. 23 (2013), the DK+ choice
 *   β = max(gₖ₊₁ᵀyₖ/dₖᵀyₖ − ‖yₖ‖²·gₖ₊₁ᵀdₖ/(dₖᵀyₖ)², η·gₖ₊₁ᵀdₖ/‖dₖ‖²), η = 0.5, yₖ = gₖ₊₁ − gₖ,
 * dₖ₊₁ = −gₖ₊₁ + β dₖ, restarting with d = −g when dᵀy ≤ 0 or d is not a descent direction.  Step lengths:
 * backtracking Armijo (the L-BFGS constants: c = 1e-4, ×0.5, at most lbfgs_ls_max trials) from a first trial of
 * the learning rate 1e-2 in the first iteration and αₖ₋₁·gₖ₋₁ᵀdₖ₋₁/gₖᵀdₖ after it (Nocedal & Wright, eq. 3.60),
 * capped at 1.
 * With a preconditioner P = L·Lᵀ (set by orc_ts_set_precond; time_step_solver_ncg_with_ext_prec,
 * stepper/nonlinear_cg.hpp:41-97) P·g replaces g in the direction and the restart, and the first term of β is
 * g₊ᵀPy/dᵀy − yᵀPy·g₊ᵀd/(dᵀy)², evaluated as aᵀ(a − b) and ‖a − b‖² with a = Lᵀg₊, b = Lᵀg (the Euclidean floor is
 * kept); with L = I every quantity equals the unpreconditioned one. */
double orc_ts_compute_step_ncg(orc_ts* ts, double* stats) {
  const int n = ts->n;
  const int pc = ts->prec != NULL;
  double* u = ts->u_next;
  double* d = (double*)malloc(sizeof(double) * n);
  double* ut = (double*)malloc(sizeof(double) * n);
  double* gt = (double*)malloc(sizeof(double) * n);
  double* pa = (double*)malloc(sizeof(double) * n); /* Lᵀg */
  double* pb = (double*)malloc(sizeof(double) * n); /* Lᵀg₊ */
  double* pg = (double*)malloc(sizeof(double) * n); /* P·g */
  int it = 0, evals = 0, converged = 0, ls_fail = 0;
  double E = orc_ts_energy_grad(ts, u, ts->g);
  evals++;
  if (pc) prec_apply(ts, ts->g, pa, pg);
  for (int i = 0; i < n; ++i) d[i] = pc ? -pg[i] : -ts->g[i];
  double alpha_prev = 1e-2, gd_prev = 0.0;
  for (; it < ts->lbfgs_maxit;) {
    const double gg = dot_n(n, ts->g, ts->g);
    if (sqrt(gg) < ts->tol_abs) {
      converged = 1;
      break;
    }
    double gd = dot_n(n, ts->g, d);
    if (!(gd < 0)) { /* not a descent direction: restart */
      for (int i = 0; i < n; ++i) d[i] = pc ? -pg[i] : -ts->g[i];
      gd = pc ? -dot_n(n, pa, pa) : -gg;
    }
    double alpha = it == 0 ? 1e-2 : alpha_prev * gd_prev / gd;
    if (!(alpha > 0) || alpha > 1.0) alpha = 1.0;
    double Et = 0;
    int ok = 0;
    for (int ls = 0; ls < ts->lbfgs_ls_max; ++ls) {
      for (int i = 0; i < n; ++i) ut[i] = u[i] + alpha * d[i];
      Et = orc_ts_energy_grad(ts, ut, gt);
      evals++;
      if (Et <= E + ts->lbfgs_c * alpha * gd) {
        ok = 1;
        break;
      }
      alpha *= ts->lbfgs_beta;
    }
    if (!ok) ls_fail++;
    if (pc) prec_apply(ts, gt, pb, pg);
    /* dᵀy, g₊ᵀPy, yᵀPy, g₊ᵀd, dᵀd with y = g₊ − g (P = I: a = g₊, b = g) */
    double dy = 0, gy = 0, yy = 0, gnd = 0, dd = 0;
    for (int i = 0; i < n; ++i) {
      const double y = gt[i] - ts->g[i];
      const double ai = pc ? pb[i] : gt[i];
      const double ya = ai - (pc ? pa[i] : ts->g[i]);
      dy += d[i] * y;
      gy += ai * ya;
      yy += ya * ya;
      gnd += gt[i] * d[i];
      dd += d[i] * d[i];
    }
    double beta = 0.0;
    if (dy > 0) {
      const double bdk = gy / dy - yy * gnd / (dy * dy);
      const double floor_ = 0.5 * gnd / dd;
      beta = bdk > floor_ ? bdk : floor_;
    }
    for (int i = 0; i < n; ++i) d[i] = -(pc ? pg[i] : gt[i]) + beta * d[i];
    memcpy(u, ut, sizeof(double) * n);
    memcpy(ts->g, gt, sizeof(double) * n);
    if (pc) {
      double* t = pa;
      pa = pb;
      pb = t;
    }
    E = Et;
    alpha_prev = alpha;
    gd_prev = gd;
    it++;
  }
  orc_ts_energy_grad(ts, u, ts->g);
  const double gn = sqrt(dot_n(n, ts->g, ts->g));
  if (stats) {
    stats[0] = it;
    stats[1] = 0;
    stats[2] = evals;
    stats[3] = converged;
    stats[4] = gn;
    stats[5] = ls_fail;
  }
  free(d);
  free(ut);
  free(gt);
  free(pa);
  free(pb);
  free(pg);
  return gn;
}

/* forward Euler stepper (time_step_solver_forward_euler, time_step.hpp:717-736): one gradient evaluation at the
 * predictor, then u_next += ½·dt²·forces, then compute_step's final evaluation. */
double orc_ts_compute_step_fe(orc_ts* ts, double* stats) {
  const int n = ts->n;
  const double h = 0.5 * ts->dt * ts->dt;
  orc_ts_energy_grad(ts, ts->u_next, ts->g);
  for (int i = 0; i < n; ++i) ts->u_next[i] += h * ts->g[i];
  orc_ts_energy_grad(ts, ts->u_next, ts->g);
  const double gn = sqrt(dot_n(n, ts->g, ts->g));
  if (stats) {
    stats[0] = 1; /* one explicit update */
    stats[1] = 0;
    stats[2] = 1; /* the solver's own evaluation; the final one is not counted, as for the other steppers */
    stats[3] = 1;
    stats[4] = gn;
    stats[5] = 0;
  }
  return gn;
}

/* step_next: u_prev ← u; u ← u_next; v = (u − u_prev)/dt, then v ← (1 − damping)·v
 * (finite_elements/time_step.hpp:530-543: copy, axpy(−1, prev), scal(1/dt), scal(1 − damping)) */
void orc_ts_step_next(orc_ts* ts) {
  for (int i = 0; i < ts->n; ++i) {
    ts->u_prev[i] = ts->u[i];
    ts->u[i] = ts->u_next[i];
    double v = ts->u[i];
    v += -1.0 * ts->u_prev[i];
    v *= 1.0 / ts->dt;
    v *= 1.0 - ts->damping;
    ts->vel[i] = v;
  }
}

/* set_damping (time_step.hpp:337-341; prepare_step checks 0 ≤ damping < 1, :499-500) */
int orc_ts_set_damping(orc_ts* ts, double damping) {
  if (!(damping >= 0.0 && damping < 1.0)) return 1;
  ts->damping = damping;
  return 0;
}

/* ---- P1 triangles in 2D: basic_time_step<..., PhysicalDim 2, TopologyDim 3>, element by element ----
 * Route of the reference: Dm⁻¹ from Dm = [x1−x0, x2−x0] (def_grad.hpp:68-84), the area |det Dm|/2
 * (rest_vol.hpp:33-50), F = Ds·Dm⁻¹ (def_grad.hpp:109-127), ∂F/∂x as the explicit 4×6 matrix of compute_pfpx_2
 * (elast/internal_physic.hpp:78-100: row 2i+j for F_ij, column 2a+c for vertex a, component c), the vertex
 * gradient from stress_element_to_vertex (time_step.hpp:62-83: P·Dm⁻ᵀ for vertices 1, 2 and minus their sum
 * for vertex 0, times A), the stiffness A·pfpxᵀ·H⁺·pfpx (time_step.hpp:41-59, the clamp through orc_sym_eig),
 * and the mass and Laplacian from the reference-triangle tables (bilinears.hpp:27-41) through the Jacobian
 * integrator (bilinears.hpp:65-91, 191-264; det of the Jacobian = 2A). */
static const double kP12ff[3][3] = {{1.0 / 12, 1.0 / 24, 1.0 / 24}, {1.0 / 24, 1.0 / 12, 1.0 / 24},
                                    {1.0 / 24, 1.0 / 24, 1.0 / 12}};
/* ∂φ_i/∂ξ on the reference triangle, φ0 = 1 − ξ − η, φ1 = ξ, φ2 = η */
static const double kP12grad[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};

void orc_tri_eval(int model, int nv, int nt, const double* X, const int* cells, const double* u, const double* lam,
                  const double* mu, double rho, int make_spsd, double* area, double* dminv, double* F,
                  double* energy, double* grad, double* stiff, double* mass, double* lap) {
  for (int e = 0; e < nt; ++e) {
    const int* c = cells + 3 * e;
    int ok = 1;
    for (int a = 0; a < 3; ++a) ok &= c[a] >= 0 && c[a] < nv;
    if (!ok) {
      const double nan = NAN;
      if (area) area[e] = nan;
      if (energy) energy[e] = nan;
      for (int k = 0; k < 4; ++k) {
        if (dminv) dminv[4 * e + k] = nan;
        if (F) F[4 * e + k] = nan;
      }
      for (int k = 0; k < 6; ++k)
        if (grad) grad[6 * e + k] = nan;
      for (int k = 0; k < 9; ++k) {
        if (mass) mass[9 * e + k] = nan;
        if (lap) lap[9 * e + k] = nan;
      }
      for (int k = 0; k < 36; ++k)
        if (stiff) stiff[36 * e + k] = nan;
      continue;
    }
    double Dm[2][2], Ds[2][2];
    for (int i = 0; i < 2; ++i)
      for (int k = 0; k < 2; ++k) {
        Dm[i][k] = X[2 * c[k + 1] + i] - X[2 * c[0] + i];
        Ds[i][k] = (X[2 * c[k + 1] + i] + u[2 * c[k + 1] + i]) - (X[2 * c[0] + i] + u[2 * c[0] + i]);
      }
    const double det = Dm[0][0] * Dm[1][1] - Dm[0][1] * Dm[1][0];
    double B[2][2]; /* Dm⁻¹ */
    B[0][0] = Dm[1][1] / det;
    B[0][1] = -Dm[0][1] / det;
    B[1][0] = -Dm[1][0] / det;
    B[1][1] = Dm[0][0] / det;
    const double A = fabs(det) / 2;
    double f[4];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) f[2 * i + j] = Ds[i][0] * B[0][j] + Ds[i][1] * B[1][j];
    double P[4], H[16];
    double psi;
    if (model == 1) {
      psi = orc_lin2_stress(f, lam[e], mu[e], P);
      orc_lin2_hessian(lam[e], mu[e], H);
    } else {
      psi = orc_snh2_stress(f, lam[e], mu[e], P);
      orc_snh2_hessian(f, lam[e], mu[e], H);
    }
    if (area) area[e] = A;
    if (energy) energy[e] = A * psi;
    for (int k = 0; k < 4; ++k) {
      if (dminv) dminv[4 * e + k] = B[k / 2][k % 2];
      if (F) F[4 * e + k] = f[k];
    }
    if (grad) {
      for (int j = 0; j < 2; ++j) {
        double tot = 0;
        for (int v = 1; v < 3; ++v) {
          /* (P·Dm⁻ᵀ)(j, v−1) = Σ_k P(j,k)·B(v−1,k) */
          const double val = P[2 * j] * B[v - 1][0] + P[2 * j + 1] * B[v - 1][1];
          grad[6 * e + 2 * v + j] = A * val;
          tot += val;
        }
        grad[6 * e + j] = -A * tot;
      }
    }
    if (stiff) {
      const double m_ = B[0][0], n_ = B[0][1], p_ = B[1][0], q_ = B[1][1];
      double pf[4][6];
      memset(pf, 0, sizeof pf);
#define PF(i, j, vert, comp, val) pf[(i) * 2 + (j)][(vert) * 2 + (comp)] = (val)
      PF(0, 0, 0, 0, -m_ - p_);
      PF(0, 0, 1, 0, m_);
      PF(0, 0, 2, 0, p_);
      PF(1, 0, 0, 1, -m_ - p_);
      PF(1, 0, 1, 1, m_);
      PF(1, 0, 2, 1, p_);
      PF(0, 1, 0, 0, -n_ - q_);
      PF(0, 1, 1, 0, n_);
      PF(0, 1, 2, 0, q_);
      PF(1, 1, 0, 1, -n_ - q_);
      PF(1, 1, 1, 1, n_);
      PF(1, 1, 2, 1, q_);
#undef PF
      double Hp[16];
      if (make_spsd) {
        double w[4], Q[16];
        orc_sym_eig(4, H, w, Q);
        for (int i = 0; i < 4; ++i) w[i] = w[i] > 1e-3 ? w[i] : 1e-3;
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 4; ++j) {
            double t = 0;
            for (int k = 0; k < 4; ++k) t += Q[i * 4 + k] * w[k] * Q[j * 4 + k];
            Hp[i * 4 + j] = t;
          }
      } else {
        memcpy(Hp, H, sizeof Hp);
      }
      for (int r = 0; r < 6; ++r)
        for (int s_ = 0; s_ < 6; ++s_) {
          double t = 0;
          for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) t += pf[i][r] * Hp[i * 4 + j] * pf[j][s_];
          stiff[36 * e + 6 * r + s_] = A * t;
        }
    }
    if (mass || lap) {
      const double detJ = fabs(det); /* p1_element_2d: det of the Jacobian [x1−x0, x2−x0] */
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          if (mass) mass[9 * e + 3 * i + j] = rho * detJ * kP12ff[i][j];
          if (lap) {
            /* Σ_k integrate_pf_pf(i, j, k, k): det·(J⁻¹ column k)ᵀ·T_ij·(J⁻¹ column k), T_ij[P][Q] = ½·∂φ_i/∂ξ_P·∂φ_j/∂ξ_Q */
            double t = 0;
            for (int k = 0; k < 2; ++k)
              for (int P_ = 0; P_ < 2; ++P_)
                for (int Q_ = 0; Q_ < 2; ++Q_)
                  t += B[P_][k] * 0.5 * kP12grad[i][P_] * kP12grad[j][Q_] * B[Q_][k];
            lap[9 * e + 3 * i + j] = detJ * t;
          }
        }
    }
  }
}
