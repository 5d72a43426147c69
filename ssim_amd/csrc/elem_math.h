// elem_math.h -- per-tetrahedron math of the FEM step (device + host inline): stable Neo-Hookean and linear
// elasticity.
//
// Conventions (as in the reference): row-major 3x3 matrices, F[a][c] = Σ_k Ds[a][k]·B[k][c] with
// Ds = [x1−x0 | x2−x0 | x3−x0] and B = Dm⁻¹ (finite_elements/def_grad.hpp:229-288); the 9x9 F-space Hessian
// is indexed 3a+c (elast/stable_neohookean.hpp:79-90); shape gradients D[0][c] = −Σ_k B[k][c],
// D[v][c] = B[v−1][c] (elast/internal_physic.hpp:106-160).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#define SSIM_HD __host__ __device__ __forceinline__

namespace ssim_hip {

template <typename T>
SSIM_HD T det3(const T A[9]) {
  return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
         A[2] * (A[3] * A[7] - A[4] * A[6]);
}

// rows of cof F are f1×f2, f2×f0, f0×f1 (elast/internal_physic.hpp:40-48)
template <typename T>
SSIM_HD void cofactor(const T F[9], T C[9]) {
  C[0] = F[4] * F[8] - F[5] * F[7];
  C[1] = F[5] * F[6] - F[3] * F[8];
  C[2] = F[3] * F[7] - F[4] * F[6];
  C[3] = F[7] * F[2] - F[8] * F[1];
  C[4] = F[8] * F[0] - F[6] * F[2];
  C[5] = F[6] * F[1] - F[7] * F[0];
  C[6] = F[1] * F[5] - F[2] * F[4];
  C[7] = F[2] * F[3] - F[0] * F[5];
  C[8] = F[0] * F[4] - F[1] * F[3];
}

// B = Dm⁻¹ (row-major) and V = |det Dm|/6: def_grad.hpp:229-247 (Eigen's 3x3 inverse, which scales the
// cofactors by one reciprocal of the determinant, as here), rest_vol.hpp:142-162
SSIM_HD void rest_shape(const double x0[3], const double x1[3], const double x2[3], const double x3[3],
                        double B[9], double* V) {
  double Dm[9];
  for (int a = 0; a < 3; ++a) {
    Dm[a * 3 + 0] = x1[a] - x0[a];
    Dm[a * 3 + 1] = x2[a] - x0[a];
    Dm[a * 3 + 2] = x3[a] - x0[a];
  }
  double C[9];
  cofactor(Dm, C);  // adj(Dm) = Cᵀ
  const double d = det3(Dm);
  const double inv = 1.0 / d;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) B[i * 3 + j] = C[j * 3 + i] * inv;
  *V = fabs(d / 6);
}

// Stable Neo-Hookean constants with δ = 1 (elast/stable_neohookean.hpp:29-35, :101-103)
template <typename T>
struct SnhScalars {
  T Ic, J, alpha, psi;
  T d2;  // μ(1 − 1/(Ic+δ))  = 2 ∂Ψ/∂Ic
  T c4;  // 2μ/(Ic+δ)²       = 4 ∂²Ψ/∂Ic²
  T s3;  // λ(J − α)         = ∂Ψ/∂J
};

template <typename T>
SSIM_HD SnhScalars<T> snh_scalars(const T F[9], T lam, T mu) {
  SnhScalars<T> s;
  const T del = T(1);
  T Ic = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) Ic += F[i] * F[i];
  s.Ic = Ic;
  s.J = det3(F);
  s.alpha = (T(1) - T(1) / (T(3) + del)) * mu / lam + T(1);
  const T Jm = s.J - s.alpha;
  const T zero_energy = T(0.5) * (lam * (T(1) - s.alpha) * (T(1) - s.alpha) - mu * log(T(3) + del));
  s.psi = T(0.5) * (mu * (Ic - T(3)) + lam * Jm * Jm - mu * log(Ic + del)) - zero_energy;
  s.d2 = mu * (T(1) - T(1) / (Ic + del));
  s.c4 = mu * (T(2) / (Ic + del) / (Ic + del));
  s.s3 = lam * Jm;
  return s;
}

// P = λ(J−α) cof F + μ(1 − 1/(Ic+δ)) F (elast/stable_neohookean.hpp:52-61)
template <typename T>
SSIM_HD void snh_stress(const T F[9], const T C[9], const SnhScalars<T>& s, T P[9]) {
#pragma unroll
  for (int i = 0; i < 9; ++i) P[i] = s.s3 * C[i] + s.d2 * F[i];
}

// Upper triangle of the 9x9 Hessian, packed row-wise: idx(i,j) for i ≤ j.
SSIM_HD constexpr int sym9(int i, int j) { return i <= j ? i * 9 - i * (i - 1) / 2 + (j - i) : j * 9 - j * (j - 1) / 2 + (i - j); }

// H = d2·I + c4·f fᵀ + λ·g gᵀ + s3·H_J with f = vec F, g = vec cof F (elast/stable_neohookean.hpp:64-91,
// add_HJ elast/internal_physic.hpp:65-76).  H_J blocks (row block a, column block b) are
// [1,0]=x̂(f2), [0,1]=−x̂(f2), [2,0]=−x̂(f1), [0,2]=x̂(f1), [2,1]=x̂(f0), [1,2]=−x̂(f0).
template <typename T>
SSIM_HD void snh_hessian_packed(const T F[9], const T C[9], T lam, const SnhScalars<T>& s, T H[45]) {
#pragma unroll
  for (int i = 0; i < 9; ++i)
#pragma unroll
    for (int j = i; j < 9; ++j) H[sym9(i, j)] = (i == j ? s.d2 : T(0)) + s.c4 * F[i] * F[j] + lam * C[i] * C[j];
  // s3·x̂(v) in block (ra, cb) with ra < cb: only the upper triangle is stored, so add the transpose of the
  // lower-block terms.  Block (0,1) = −x̂(f2), (0,2) = +x̂(f1), (1,2) = −x̂(f0) (the lower blocks are their
  // transposes, x̂ being antisymmetric).
  const T s3 = s.s3;
  const T* f0 = F;
  const T* f1 = F + 3;
  const T* f2 = F + 6;
  auto add_block = [&](int ra, int cb, const T* v, T sc) {
    // x̂(v) = [[0,−v2,v1],[v2,0,−v0],[−v1,v0,0]]
    const T xh[9] = {T(0), -v[2], v[1], v[2], T(0), -v[0], -v[1], v[0], T(0)};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) H[sym9(3 * ra + i, 3 * cb + j)] += sc * xh[i * 3 + j];
  };
  add_block(0, 1, f2, -s3);
  add_block(0, 2, f1, s3);
  add_block(1, 2, f0, -s3);
}

// Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix (Golub & Van Loan, Alg. 8.5.3).
// On return A holds the eigenvalues on its diagonal and V (row-major, columns = eigenvectors) the basis.
template <int p, int q, typename T>
SSIM_HD void jacobi_rot3(T A[3][3], T V[3][3]) {
  constexpr int r = 3 - p - q;
  const T apq = A[p][q];
  if (apq == T(0)) return;
  const T theta = (A[q][q] - A[p][p]) / (T(2) * apq);
  const T at = fabs(theta);
  T t;
  if (at > T(1e30))
    t = T(0.5) / theta;
  else
    t = (theta >= T(0) ? T(1) : T(-1)) / (at + sqrt(theta * theta + T(1)));
  const T c = T(1) / sqrt(t * t + T(1));
  const T s = t * c;
  const T arp = A[r][p], arq = A[r][q];
  A[r][p] = A[p][r] = c * arp - s * arq;
  A[r][q] = A[q][r] = s * arp + c * arq;
  A[p][p] -= t * apq;
  A[q][q] += t * apq;
  A[p][q] = A[q][p] = T(0);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const T vkp = V[k][p], vkq = V[k][q];
    V[k][p] = c * vkp - s * vkq;
    V[k][q] = s * vkp + c * vkq;
  }
}

template <typename T>
SSIM_HD void eig3_jacobi(T A[3][3], T V[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) V[i][j] = (i == j) ? T(1) : T(0);
  const T scale = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2] +
                  T(2) * (A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2]);
  const T eps2 = sizeof(T) == 8 ? T(1e-34) : T(1e-16);
  for (int sweep = 0; sweep < 12; ++sweep) {
    const T off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    if (!(off > eps2 * scale)) break;
    jacobi_rot3<0, 1>(A, V);
    jacobi_rot3<0, 2>(A, V);
    jacobi_rot3<1, 2>(A, V);
  }
}

// Rotation-variant SVD F = U diag(σ) Vᵀ with U, V ∈ SO(3) and σ0 ≥ σ1 ≥ |σ2| (σ2 < 0 iff det F < 0).
// V from the eigenvectors of FᵀF (Jacobi), U by Gram–Schmidt on the columns of F·V, σ_i = u_i·(F v_i).
// U and V are row-major with the singular vectors as columns.
template <typename T>
SSIM_HD void svd3(const T F[9], T U[9], T sig[3], T Vm[9]) {
  T A[3][3], V[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) A[i][j] = F[0 * 3 + i] * F[0 * 3 + j] + F[1 * 3 + i] * F[1 * 3 + j] + F[2 * 3 + i] * F[2 * 3 + j];
  eig3_jacobi(A, V);
  T ev[3] = {A[0][0], A[1][1], A[2][2]};
  // sort descending (swap columns of V)
  auto swapc = [&](int a, int b) {
    T t = ev[a];
    ev[a] = ev[b];
    ev[b] = t;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      T x = V[k][a];
      V[k][a] = V[k][b];
      V[k][b] = x;
    }
  };
  if (ev[0] < ev[1]) swapc(0, 1);
  if (ev[0] < ev[2]) swapc(0, 2);
  if (ev[1] < ev[2]) swapc(1, 2);
  // make V a proper rotation
  const T dv = V[0][0] * (V[1][1] * V[2][2] - V[1][2] * V[2][1]) - V[0][1] * (V[1][0] * V[2][2] - V[1][2] * V[2][0]) +
               V[0][2] * (V[1][0] * V[2][1] - V[1][1] * V[2][0]);
  if (dv < T(0)) {
#pragma unroll
    for (int k = 0; k < 3; ++k) V[k][2] = -V[k][2];
  }
  // b_i = F v_i
  T b[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int a = 0; a < 3; ++a) b[i][a] = F[a * 3 + 0] * V[0][i] + F[a * 3 + 1] * V[1][i] + F[a * 3 + 2] * V[2][i];
  T u0[3], u1[3], u2[3];
  T n0 = sqrt(b[0][0] * b[0][0] + b[0][1] * b[0][1] + b[0][2] * b[0][2]);
  if (n0 > T(0)) {
    for (int a = 0; a < 3; ++a) u0[a] = b[0][a] / n0;
  } else {
    u0[0] = T(1); u0[1] = T(0); u0[2] = T(0);
  }
  const T p01 = u0[0] * b[1][0] + u0[1] * b[1][1] + u0[2] * b[1][2];
  T w1[3] = {b[1][0] - p01 * u0[0], b[1][1] - p01 * u0[1], b[1][2] - p01 * u0[2]};
  T n1 = sqrt(w1[0] * w1[0] + w1[1] * w1[1] + w1[2] * w1[2]);
  const T tiny = sizeof(T) == 8 ? T(1e-13) : T(1e-6);
  if (!(n1 > tiny * (n0 > T(0) ? n0 : T(1)))) {
    // rank ≤ 1: any unit vector orthogonal to u0
    const int m = (fabs(u0[0]) <= fabs(u0[1]) && fabs(u0[0]) <= fabs(u0[2])) ? 0 : (fabs(u0[1]) <= fabs(u0[2]) ? 1 : 2);
    T e[3] = {T(0), T(0), T(0)};
    e[m] = T(1);
    const T pe = u0[m];
    w1[0] = e[0] - pe * u0[0];
    w1[1] = e[1] - pe * u0[1];
    w1[2] = e[2] - pe * u0[2];
    n1 = sqrt(w1[0] * w1[0] + w1[1] * w1[1] + w1[2] * w1[2]);
  }
  for (int a = 0; a < 3; ++a) u1[a] = w1[a] / n1;
  u2[0] = u0[1] * u1[2] - u0[2] * u1[1];
  u2[1] = u0[2] * u1[0] - u0[0] * u1[2];
  u2[2] = u0[0] * u1[1] - u0[1] * u1[0];
  sig[0] = u0[0] * b[0][0] + u0[1] * b[0][1] + u0[2] * b[0][2];
  sig[1] = u1[0] * b[1][0] + u1[1] * b[1][1] + u1[2] * b[1][2];
  sig[2] = u2[0] * b[2][0] + u2[1] * b[2][1] + u2[2] * b[2][2];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    U[a * 3 + 0] = u0[a];
    U[a * 3 + 1] = u1[a];
    U[a * 3 + 2] = u2[a];
#pragma unroll
    for (int i = 0; i < 3; ++i) Vm[a * 3 + i] = V[a][i];
  }
}

// H⁺ = Q·max(Λ, floor)·Qᵀ applied as H + Σ_{λ_i < floor} (floor − λ_i) q_i q_iᵀ, using the analytic eigensystem
// of an energy Ψ(Ic, J) in the SVD frame F = U Σ Vᵀ (Smith, de Goes, Kim, "Analytic Eigensystems for
// Isotropic Distortion Energies", ACM TOG 38(1), 2019):
//   flip  q = vec(u_i v_jᵀ + u_j v_iᵀ)/√2, λ = d2 − s3·σ_k
//   twist q = vec(u_i v_jᵀ − u_j v_iᵀ)/√2, λ = d2 + s3·σ_k        ({i,j,k} = {0,1,2})
//   scaling q = Σ_i w_i vec(u_i v_iᵀ) with (w, λ) the eigenpairs of
//           A = d2·I + c4·σσᵀ + λ·ĝĝᵀ + s3·[[0,σ2,σ1],[σ2,0,σ0],[σ1,σ0,0]],  ĝ = (σ1σ2, σ0σ2, σ0σ1).
// The floor is the reference's absolute 1e-3 on the unscaled 9x9 (finite_elements/time_step.hpp:50-56).
template <typename T>
SSIM_HD void snh_project_spd(const T F[9], T lam, const SnhScalars<T>& s, T floor_, T H[45]) {
  T U[9], sg[3], V[9];
  svd3(F, U, sg, V);
  auto rank1 = [&](const T Q[9], T coef) {
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
      for (int j = i; j < 9; ++j) H[sym9(i, j)] += coef * Q[i] * Q[j];
  };
  const T r2 = T(0.70710678118654752440);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int i = (k + 1) % 3, j = (k + 2) % 3;
    const T lf = s.d2 - s.s3 * sg[k];
    const T lt = s.d2 + s.s3 * sg[k];
    if (lf < floor_ || lt < floor_) {
      T Qf[9], Qt[9];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const T x = U[a * 3 + i] * V[c * 3 + j], y = U[a * 3 + j] * V[c * 3 + i];
          Qf[a * 3 + c] = (x + y) * r2;
          Qt[a * 3 + c] = (x - y) * r2;
        }
      if (lf < floor_) rank1(Qf, floor_ - lf);
      if (lt < floor_) rank1(Qt, floor_ - lt);
    }
  }
  T A[3][3], W[3][3];
  const T g[3] = {sg[1] * sg[2], sg[0] * sg[2], sg[0] * sg[1]};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) A[i][j] = (i == j ? s.d2 : T(0)) + s.c4 * sg[i] * sg[j] + lam * g[i] * g[j];
  A[0][1] += s.s3 * sg[2];
  A[1][0] += s.s3 * sg[2];
  A[0][2] += s.s3 * sg[1];
  A[2][0] += s.s3 * sg[1];
  A[1][2] += s.s3 * sg[0];
  A[2][1] += s.s3 * sg[0];
  eig3_jacobi(A, W);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const T lm = A[m][m];
    if (lm < floor_) {
      T Q[9];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          Q[a * 3 + c] = W[0][m] * U[a * 3 + 0] * V[c * 3 + 0] + W[1][m] * U[a * 3 + 1] * V[c * 3 + 1] +
                         W[2][m] * U[a * 3 + 2] * V[c * 3 + 2];
      rank1(Q, floor_ - lm);
    }
  }
}

// ------------------------------------------------------------------------------------------------------
// Linear elasticity (elast/linear.hpp:24-69; strain elast/internal_physic.hpp:16-19):
//   ε = ½(F + Fᵀ) − I,  Ψ = μ‖ε‖² + ½λ·tr(ε)²,  P = 2μ·ε + λ·tr(ε)·I,
//   H = μ·(I + S) + λ·t tᵀ with S the transposition vec(F) → vec(Fᵀ) and t = vec(I).

template <typename T>
SSIM_HD T linear_stress(const T F[9], T lam, T mu, T P[9]) {
  const T tr = F[0] + F[4] + F[8] - T(3);
  T n2 = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T e = T(0.5) * (F[3 * a + c] + F[3 * c + a]) - (a == c ? T(1) : T(0));
      n2 += e * e;
      P[3 * a + c] = T(2) * mu * e + (a == c ? lam * tr : T(0));
    }
  return mu * n2 + T(0.5) * lam * tr * tr;
}

template <typename T>
SSIM_HD void linear_hessian_packed(T lam, T mu, T H[45]) {
#pragma unroll
  for (int i = 0; i < 9; ++i)
#pragma unroll
    for (int j = i; j < 9; ++j) {
      const int a = i / 3, b = i % 3, c = j / 3, d = j % 3;
      H[sym9(i, j)] = (i == j ? mu : T(0)) + (a == d && b == c ? mu : T(0)) + (a == b && c == d ? lam : T(0));
    }
}

// SPD projection in closed form.  H has three eigenspaces: the skew matrices (eigenvalue 0, projector ½(I − S)),
// the trace-free symmetric matrices (2μ, projector ½(I + S) − t tᵀ/3) and the multiples of I (2μ + 3λ,
// projector t tᵀ/3).  Eigenvalues under the floor are raised to it, as the 9x9 eigen-decomposition of the
// reference does (finite_elements/time_step.hpp:50-56).
template <typename T>
SSIM_HD void linear_project_spd(T lam, T mu, T floor_, T H[45]) {
  const T ka = floor_ > T(0) ? floor_ : T(0);
  const T kd = floor_ > T(2) * mu ? floor_ - T(2) * mu : T(0);
  const T kv = floor_ > T(2) * mu + T(3) * lam ? floor_ - T(2) * mu - T(3) * lam : T(0);
#pragma unroll
  for (int i = 0; i < 9; ++i)
#pragma unroll
    for (int j = i; j < 9; ++j) {
      const int a = i / 3, b = i % 3, c = j / 3, d = j % 3;
      const T id = i == j ? T(0.5) : T(0);
      const T sw = a == d && b == c ? T(0.5) : T(0);
      const T tt = a == b && c == d ? T(1) / T(3) : T(0);
      H[sym9(i, j)] += ka * (id - sw) + kd * (id + sw - tt) + kv * tt;
    }
}

// ------------------------------------------------------------------------------------------------------
// Material dispatch: the element bodies call these two with the model as a template argument.

enum MaterialModel : int { kStableNeoHookean = 0, kLinearElastic = 1 };

template <int Model, typename T>
SSIM_HD T material_stress(const T F[9], T lam, T mu, T P[9]) {
  if constexpr (Model == kLinearElastic) {
    return linear_stress(F, lam, mu, P);
  } else {
    const SnhScalars<T> s = snh_scalars(F, lam, mu);
    T C[9];
    cofactor(F, C);
    snh_stress(F, C, s, P);
    return s.psi;
  }
}

template <int Model, typename T>
SSIM_HD void material_hessian(const T F[9], T lam, T mu, int make_spsd, T floor_, T H[45]) {
  if constexpr (Model == kLinearElastic) {
    linear_hessian_packed(lam, mu, H);
    if (make_spsd) linear_project_spd(lam, mu, floor_, H);
  } else {
    const SnhScalars<T> s = snh_scalars(F, lam, mu);
    T C[9];
    cofactor(F, C);
    snh_hessian_packed(F, C, lam, s, H);
    if (make_spsd) snh_project_spd(F, lam, s, floor_, H);
  }
}

// ------------------------------------------------------------------------------------------------------
// 2D models (Ndim = 2 of elast/stable_neohookean.hpp and elast/linear.hpp; the reference has no 2D time step,
// they are exercised by element tests, tests/fem_intro/main.cpp:155-300).  F row-major [F00, F01, F10, F11];
// the 4x4 Hessian is full, H[4·i + j] = ∂²Ψ/∂F_i∂F_j in that flattening.

// δ of the 2D stable Neo-Hookean model: the positive root of (2t+1)δ² + (2t+2)δ − 2t = 0, t = μ/λ
// (elast/stable_neohookean.hpp:96-107; δ = 1 in 3D)
template <typename T>
SSIM_HD T snh2_delta(T lam, T mu) {
  const T t = mu / lam;
  const T a = T(2) * t + T(1), b = T(2) * t + T(2), c = -T(2) * t;
  return (-b + sqrt(b * b - T(4) * a * c)) / (T(2) * a);
}

// Ψ = ½(μ(Ic − 2) + λ(J − α)² − μ·ln(Ic + δ)) − Ψ(I), α = (1 − 1/(2+δ))·μ/λ + 1, P = λ(J − α)·cof F +
// μ(1 − 1/(Ic+δ))·F with cof F = [[F11, −F10], [−F01, F00]] (elast/stable_neohookean.hpp:23-62, 2D
// partial_determinant elast/internal_physic.hpp:32-37)
template <typename T>
SSIM_HD T snh2_stress(const T F[4], T lam, T mu, T P[4]) {
  const T del = snh2_delta(lam, mu);
  const T Ic = F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + F[3] * F[3];
  const T J = F[0] * F[3] - F[1] * F[2];
  const T alpha = (T(1) - T(1) / (T(2) + del)) * mu / lam + T(1);
  const T Jm = J - alpha;
  const T C[4] = {F[3], -F[2], -F[1], F[0]};
  const T d2 = mu * (T(1) - T(1) / (Ic + del));
#pragma unroll
  for (int i = 0; i < 4; ++i) P[i] = lam * Jm * C[i] + d2 * F[i];
  const T zero_energy = T(0.5) * (lam * (T(1) - alpha) * (T(1) - alpha) - mu * log(T(2) + del));
  return T(0.5) * (mu * (Ic - T(2)) + lam * Jm * Jm - mu * log(Ic + del)) - zero_energy;
}

// H = μ(1 − 1/(Ic+δ))·I + 2μ/(Ic+δ)²·f fᵀ + λ·g gᵀ + λ(J − α)·H_J, f = vec F, g = vec cof F, H_J = ∂²J/∂F²
// (entries (F00, F11) = 1, (F01, F10) = −1; add_HJ, elast/internal_physic.hpp:54-60)
template <typename T>
SSIM_HD void snh2_hessian(const T F[4], T lam, T mu, T H[16]) {
  const T del = snh2_delta(lam, mu);
  const T Ic = F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + F[3] * F[3];
  const T J = F[0] * F[3] - F[1] * F[2];
  const T alpha = (T(1) - T(1) / (T(2) + del)) * mu / lam + T(1);
  const T C[4] = {F[3], -F[2], -F[1], F[0]};
  const T d2 = mu * (T(1) - T(1) / (Ic + del)), c4 = mu * (T(2) / (Ic + del) / (Ic + del)), s3 = lam * (J - alpha);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) H[4 * i + j] = (i == j ? d2 : T(0)) + c4 * F[i] * F[j] + lam * C[i] * C[j];
  H[0 * 4 + 3] += s3;
  H[3 * 4 + 0] += s3;
  H[1 * 4 + 2] -= s3;
  H[2 * 4 + 1] -= s3;
}

template <typename T>
SSIM_HD T linear2_stress(const T F[4], T lam, T mu, T P[4]) {
  const T tr = F[0] + F[3] - T(2);
  T n2 = 0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const T e = T(0.5) * (F[2 * a + c] + F[2 * c + a]) - (a == c ? T(1) : T(0));
      n2 += e * e;
      P[2 * a + c] = T(2) * mu * e + (a == c ? lam * tr : T(0));
    }
  return mu * n2 + T(0.5) * lam * tr * tr;
}

template <typename T>
SSIM_HD void linear2_hessian(T lam, T mu, T H[16]) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int a = i / 2, b = i % 2, c = j / 2, d = j % 2;
      H[4 * i + j] = (i == j ? mu : T(0)) + (a == d && b == c ? mu : T(0)) + (a == b && c == d ? lam : T(0));
    }
}

// The SPD projection of the 2D stable Neo-Hookean Hessian in closed form.  H = d2·I + c4·f fᵀ + λ·g gᵀ + s3·H_J
// with g = vec cof F = H_J·f, and H_J has the fixed eigenbasis b1 = (e0+e3)/√2, b2 = (e1−e2)/√2 (eigenvalue +1),
// b3 = (e0−e3)/√2, b4 = (e1+e2)/√2 (−1).  In that basis f = (f₊, f₋) and g = (f₊, −f₋), so (p, 0) with p ⊥ f₊ is an
// eigenvector for d2 + s3, (0, q) with q ⊥ f₋ one for d2 − s3, and the two left, spanned by (f̂₊, 0) and (0, f̂₋),
// diagonalise the 2x2 block [[d2 + s3 + (c4+λ)|f₊|², (c4−λ)|f₊||f₋|], [·, d2 − s3 + (c4+λ)|f₋|²]] with one Jacobi
// rotation.  H⁺ = Σ max(λ_i, floor)·w_i w_iᵀ, the matrix time_step.hpp:50-56 forms with an eigensolver.
template <typename T>
SSIM_HD void snh2_project_spd(const T F[4], T lam, T mu, T floor_, T H[16]) {
  const T del = snh2_delta(lam, mu);
  const T Ic = F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + F[3] * F[3];
  const T J = F[0] * F[3] - F[1] * F[2];
  const T alpha = (T(1) - T(1) / (T(2) + del)) * mu / lam + T(1);
  const T d2 = mu * (T(1) - T(1) / (Ic + del)), c4 = mu * (T(2) / (Ic + del) / (Ic + del)), s3 = lam * (J - alpha);
  const T r = T(0.70710678118654752440);  // 1/√2
  const T fp[2] = {r * (F[0] + F[3]), r * (F[1] - F[2])}, fm[2] = {r * (F[0] - F[3]), r * (F[1] + F[2])};
  const T np_ = sqrt(fp[0] * fp[0] + fp[1] * fp[1]), nm = sqrt(fm[0] * fm[0] + fm[1] * fm[1]);
  // unit directions (any unit vector where the component vanishes: the block is then already diagonal)
  const T up[2] = {np_ > T(0) ? fp[0] / np_ : T(1), np_ > T(0) ? fp[1] / np_ : T(0)};
  const T um[2] = {nm > T(0) ? fm[0] / nm : T(1), nm > T(0) ? fm[1] / nm : T(0)};
  const T a = d2 + s3 + (c4 + lam) * np_ * np_, c = d2 - s3 + (c4 + lam) * nm * nm, b = (c4 - lam) * np_ * nm;
  T t = T(0);
  if (b != T(0)) {
    const T theta = (c - a) / (T(2) * b);
    t = fabs(theta) > T(1e150) ? T(0.5) / theta
                               : (theta >= T(0) ? T(1) : T(-1)) / (fabs(theta) + sqrt(theta * theta + T(1)));
  }
  const T cs = T(1) / sqrt(t * t + T(1)), sn = t * cs;
  // eigenvectors in the coordinates (b1, b2, b3, b4) and their eigenvalues
  const T V[4][4] = {{-up[1], up[0], T(0), T(0)},
                     {T(0), T(0), -um[1], um[0]},
                     {cs * up[0], cs * up[1], -sn * um[0], -sn * um[1]},
                     {sn * up[0], sn * up[1], cs * um[0], cs * um[1]}};
  const T w[4] = {d2 + s3, d2 - s3, a - t * b, c + t * b};
#pragma unroll
  for (int k = 0; k < 16; ++k) H[k] = T(0);
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const T lw = w[n] > floor_ ? w[n] : floor_;
    // back to F-space coordinates: Σ ṽ_i b_i
    const T x[4] = {r * (V[n][0] + V[n][2]), r * (V[n][1] + V[n][3]), r * (V[n][3] - V[n][1]),
                    r * (V[n][0] - V[n][2])};
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) H[4 * i + j] += lw * x[i] * x[j];
  }
}

// Cyclic Jacobi eigen-decomposition of a symmetric 4x4 matrix (Golub & Van Loan, Matrix Computations, 4th ed.,
// §8.5.2-8.5.3: symmetric 2x2 Schur rotations swept row by row until the off-diagonal mass is below 1e-32 of the
// total, at most 20 sweeps), then H⁺ = Q·max(Λ, floor)·Qᵀ: the SPD projection of the 2D element Hessian
// (time_step.hpp:50-56, SelfAdjointEigenSolver with eigenvalues clamped from below at 1e-3).
template <typename T>
SSIM_HD void sym4_project_spd(T H[16], T floor_) {
  T A[16], Q[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    A[k] = H[k];
    Q[k] = (k % 5 == 0) ? T(1) : T(0);
  }
  for (int sweep = 0; sweep < 20; ++sweep) {
    T off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        tot += A[4 * i + j] * A[4 * i + j];
        if (i != j) off += A[4 * i + j] * A[4 * i + j];
      }
    if (!(off > T(1e-32) * tot)) break;
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = p + 1; q < 4; ++q) {
        const T apq = A[4 * p + q];
        if (apq == T(0)) continue;
        const T tau = (A[4 * q + q] - A[4 * p + p]) / (T(2) * apq);
        const T t = (tau >= T(0) ? T(1) : T(-1)) / (fabs(tau) + sqrt(T(1) + tau * tau));
        const T c = T(1) / sqrt(T(1) + t * t), s = t * c;
        // A <- Jᵀ A J with J the rotation in the (p, q) plane; columns first, then rows
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const T akp = A[4 * k + p], akq = A[4 * k + q];
          A[4 * k + p] = c * akp - s * akq;
          A[4 * k + q] = s * akp + c * akq;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const T apk = A[4 * p + k], aqk = A[4 * q + k];
          A[4 * p + k] = c * apk - s * aqk;
          A[4 * q + k] = s * apk + c * aqk;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const T qkp = Q[4 * k + p], qkq = Q[4 * k + q];
          Q[4 * k + p] = c * qkp - s * qkq;
          Q[4 * k + q] = s * qkp + c * qkq;
        }
      }
  }
  T w[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) w[k] = A[5 * k] > floor_ ? A[5 * k] : floor_;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = i; j < 4; ++j) {
      T v = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) v += Q[4 * i + k] * w[k] * Q[4 * j + k];
      H[4 * i + j] = v;
      H[4 * j + i] = v;
    }
}

}  // namespace ssim_hip
