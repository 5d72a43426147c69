// kernels.hip -- HIP kernels of the FEM soft-body time step for gfx950 (MI355X).
//
// The path is memory-bound gather/scatter work, so no MFMA: kernels are laid out for coalesced HBM access
// (structure-of-arrays per-tet data, blocked-ELL matrix with one row per lane), deterministic reductions
// (wavefront shuffles, a fixed-order workgroup tree and a last-workgroup final pass) and atomic-free
// assembly (every element writes its contributions to precomputed, target-sorted slots; a segmented reduce
// sums each target's contiguous run).
#include <climits>

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "device_common.h"

namespace ssim_hip {

int max_partials(int n) { return (n + kMinTile - 1) / kMinTile + 1; }

__host__ __device__ static inline int nblocks(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// ------------------------------------------------------------------------------------------------------
// rest shape

// With an event pair the launch goes through hipExtLaunchKernelGGL, which stamps the events at the kernel's own
// start and end (the interval a profiler reports) instead of around the stream commands.
template <typename K, typename... Args>
void launch_timed(K kernel, int grid, int block, hipStream_t s, hipEvent_t e0, hipEvent_t e1, Args... args) {
  if (e0)
    hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, s, e0, e1, 0, args...);
  else
    kernel<<<grid, block, 0, s>>>(args...);
}

__global__ __launch_bounds__(kBlock) void k_rest_shape(const double* __restrict__ X, const int4* __restrict__ cells,
                                                       int nt, double* __restrict__ B, double* __restrict__ V) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= nt) return;
  const int4 c = cells[e];
  if (c.x < 0) {  // absent tet (padding of a batched layout): no volume, never gathered
#pragma unroll
    for (int k = 0; k < 9; ++k) B[(size_t)k * nt + e] = 0.0;
    V[e] = 0.0;
    return;
  }
  const int id[4] = {c.x, c.y, c.z, c.w};
  double x[4][3];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int a = 0; a < 3; ++a) x[i][a] = X[3 * (size_t)id[i] + a];
  double Bl[9], Vl;
  rest_shape(x[0], x[1], x[2], x[3], Bl, &Vl);
#pragma unroll
  for (int k = 0; k < 9; ++k) B[(size_t)k * nt + e] = Bl[k];
  V[e] = Vl;
}

void launch_rest_shape(const double* X, const int4* cells, int nt, double* B, double* V, hipStream_t s) {
  if (nt <= 0) return;
  k_rest_shape<<<nblocks(nt), kBlock, 0, s>>>(X, cells, nt, B, V);
}

// ------------------------------------------------------------------------------------------------------
// energy + gradient (update_energy_and_gradients, time_step.hpp:577-624), element-wise:
//   e = ½·(ρV/20)·(Σ|Δ_i|² + |ΣΔ_i|²) + dt²·V·Ψ,   g_i = (ρV/20)·(Δ_i + ΣΔ_j) + dt²·V·f_i,   Δ = u − û.
// The consistent mass M_e = ρV(1+δ_ij)/20 (bilinears.hpp:140-148) is applied element by element instead of
// through a CSR SpMV.

// Energy + gradient, tile-reduced: workgroup b takes the TILE consecutive tets of tile b (one per thread; TILE =
// 64, 128 or 256, a launch knob), writes each tet's 4 vertex contributions into LDS at the tile-local slots tslot
// (the tile's corners sorted by (vertex, tet, corner)), and sums every tile vertex's contiguous run in order into
// one partial per (tile, vertex), gpart[gt_off[b] + j].  Only those partials reach HBM (about 0.6 per tet on a
// structured bar with 256-tet tiles instead of 4 contributions); k_grad_gather adds each vertex's partials in tile
// order.
template <typename T, int Model, int TILE>
__global__ __launch_bounds__(TILE) void k_elem_grad_tile(MeshView m, const double* __restrict__ u,
                                                         const double* __restrict__ uhat, double dt2,
                                                         const ushort4* __restrict__ tslot,
                                                         const int* __restrict__ gt_off,
                                                         const int* __restrict__ gt_run,
                                                         double* __restrict__ gpart, double* __restrict__ part_e) {
  const int nt = m.nt > 0 ? m.nt : 1;
  const int lb = logical_block((nt + TILE - 1) / TILE, m.xcd_remap);
  if (lb < 0) return;  // uniform over the workgroup
  const int e = lb * TILE + threadIdx.x;
  double energy[1] = {elem_grad_tile<T, Model, TILE>(m, lb, e < m.nt && m.cells[e].x >= 0, u, uhat, dt2, tslot,
                                                     gt_off, gt_run, gpart)};
  store_partials<1, TILE>(energy, part_e, lb);
}

int launch_elem_grad_tile(const MeshView& m, int tile, const double* u, const double* uhat, double dt2,
                          const ushort4* tslot, const int* gt_off, const int* gt_run, double* gpart, double* part_e,
                          bool fp32, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  const int nt = m.nt > 0 ? m.nt : 1;
  const int nb = (nt + tile - 1) / tile;
  dispatch_elem(m.model, fp32, [&](auto k) {
    using K = decltype(k);
    const int grid = grid_for(nb, m.xcd_remap);
    if (tile == 64)
      launch_timed(k_elem_grad_tile<typename K::type, K::model, 64>, grid, 64, s, e0, e1, m, u, uhat, dt2, tslot,
                   gt_off, gt_run, gpart, part_e);
    else if (tile == 128)
      launch_timed(k_elem_grad_tile<typename K::type, K::model, 128>, grid, 128, s, e0, e1, m, u, uhat, dt2, tslot,
                   gt_off, gt_run, gpart, part_e);
    else
      launch_timed(k_elem_grad_tile<typename K::type, K::model, kBlock>, grid, kBlock, s, e0, e1, m, u, uhat, dt2,
                   tslot, gt_off, gt_run, gpart, part_e);
  });
  return nb;
}

// g[v] = Σ over the tiles that touch v (ascending) of their partials for v; Dirichlet dofs zeroed
// (boundary.hpp:45-51); partials of ‖g‖²
__global__ __launch_bounds__(kBlock) void k_grad_gather(int nv, const int* __restrict__ gv_seg,
                                                        const int* __restrict__ gv_pos,
                                                        const double* __restrict__ gpart,
                                                        const unsigned char* __restrict__ dof, double* __restrict__ g,
                                                        double* __restrict__ part_g) {
  const int v = blockIdx.x * kBlock + threadIdx.x;
  double gg[1] = {v < nv ? grad_gather_vertex(v, gv_seg, gv_pos, gpart, dof, g) : 0.0};
  store_partials<1>(gg, part_g);
}

int launch_grad_gather(int nv, const int* gv_seg, const int* gv_pos, const double* gpart, const unsigned char* dof,
                       double* g, double* part_g, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  const int nb = nblocks(nv > 0 ? nv : 1);
  launch_timed(k_grad_gather, nb, kBlock, s, e0, e1, nv, gv_seg, gv_pos, gpart, dof, g, part_g);
  return nb;
}

// Projective-dynamics matrix (stepper/projective_dynamics.hpp:51-87): the 10 oriented blocks of tet e,
// (ρV(1+δ_vw)/20 + c·V·∇N_v·∇N_w)·I₃ with c = dt²κ (consistent mass plus the scaled P1 Laplacian on every dof),
// to the Hessian's contribution slots hperm[10e+k]; the Hessian reduce then assembles them.
__global__ __launch_bounds__(kBlock) void k_elem_pd(MeshView m, double c, const int* __restrict__ hperm,
                                                    double* __restrict__ hcontrib) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= m.nt || m.cells[e].x < 0) return;
  const TetLocal t = load_tet(m, e);
  double D[4][3];
  shape_grads(t.B, D);
  const double mm = m.rho * t.V / 20.0;
  const int* pp = hperm + 10 * (size_t)e;
#pragma unroll
  for (int w = 0; w < 4; ++w)
#pragma unroll
    for (int v = 0; v <= w; ++v) {
      const double val = (v == w ? 2.0 * mm : mm) + c * t.V * (D[v][0] * D[w][0] + D[v][1] * D[w][1] + D[v][2] * D[w][2]);
      const size_t pos = 9 * (size_t)pp[pair_index(v, w)];
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int d = 0; d < 3; ++d) hcontrib[pos + 3 * b + d] = b == d ? val : 0.0;
    }
}

void launch_elem_pd(const MeshView& m, double c, const int* hperm, double* hcontrib, hipStream_t s) {
  if (m.nt <= 0) return;
  k_elem_pd<<<nblocks(m.nt), kBlock, 0, s>>>(m, c, hperm, hcontrib);
}

// segmented reduce of the vertex-sorted gradient contributions, staged through LDS; Dirichlet dofs zeroed
// (boundary.hpp:45-51)
__global__ __launch_bounds__(kBlock) void k_grad_reduce(int nv, const int* __restrict__ gseg,
                                                        const double* __restrict__ gcontrib,
                                                        const unsigned char* __restrict__ dof, double* __restrict__ g,
                                                        double* __restrict__ part_g) {
  const int v0 = blockIdx.x * kBlock;
  double gg[1] = {grad_reduce_vertices(v0, min(v0 + kBlock, nv), gseg, gcontrib, dof, g)};
  store_partials<1>(gg, part_g);
}

int launch_grad_reduce(int nv, const int* gseg, const double* gcontrib, const unsigned char* dof, double* g,
                       double* part_g, hipStream_t s) {
  const int nb = nblocks(nv > 0 ? nv : 1);
  k_grad_reduce<<<nb, kBlock, 0, s>>>(nv, gseg, gcontrib, dof, g, part_g);
  return nb;
}

// one workgroup: out[k] = Σ partials (width K) for up to two partial arrays
__global__ __launch_bounds__(kBlock) void k_finalize(const double* __restrict__ pa, int na, int ka,
                                                     const double* __restrict__ pb, int nb, double* __restrict__ out) {
  double t[2];
  if (ka == 2) {
    sum_partials<2>(pa, na, t);
  } else {
    double t1[1];
    sum_partials<1>(pa, na, t1);
    t[0] = t1[0];
    t[1] = 0;
  }
  double u1[1] = {0.0};
  if (pb) sum_partials<1>(pb, nb, u1);
  if (threadIdx.x == 0) {
    out[0] = t[0];
    if (ka == 2) out[1] = t[1];
    if (pb) out[ka] = u1[0];
  }
}

void launch_finalize(const double* pa, int na, int ka, const double* pb, int nb, double* out, hipStream_t s) {
  k_finalize<<<1, kBlock, 0, s>>>(pa, na, ka, pb, nb, out);
}

// ------------------------------------------------------------------------------------------------------
// Hessian (update_hessian, time_step.hpp:626-658; wrapped_pfpx :41-59)

template <typename T, int Model>
__global__ __launch_bounds__(kBlock) void k_elem_hess(MeshView m, const double* __restrict__ u, double dt2,
                                                      int make_spsd, const int* __restrict__ hperm,
                                                      double* __restrict__ hcontrib) {
  const int lb = logical_block(nblocks(m.nt), m.xcd_remap);
  const int e = lb * kBlock + threadIdx.x;
  if (lb < 0 || e >= m.nt) return;
  elem_hess_tet<T, Model>(m, e, u, dt2, make_spsd, hperm, hcontrib);
}

void launch_elem_hess(const MeshView& m, const double* u, double dt2, int make_spsd, const int* hperm,
                      double* hcontrib, bool fp32, hipStream_t s) {
  if (m.nt <= 0) return;
  dispatch_elem(m.model, fp32, [&](auto k) {
    using K = decltype(k);
    k_elem_hess<typename K::type, K::model>
        <<<grid_for(nblocks(m.nt), m.xcd_remap), kBlock, 0, s>>>(m, u, dt2, make_spsd, hperm, hcontrib);
  });
}

// One thread per upper block (r ≤ c): sum its contiguous contribution run, apply the Dirichlet identity
// filter (identity_hessian_impl, boundary.hpp:53-67), write block (r,c) into the symmetric blocked-ELL values
// (the lower block (c,r) is its transpose and is not stored), and the Jacobi inverse diagonal for r == c.
__global__ __launch_bounds__(kBlock) void k_hess_reduce(int nub, const int* __restrict__ hseg,
                                                        const double* __restrict__ hcontrib,
                                                        const int2* __restrict__ ub_rc, const int* __restrict__ ub_pos,
                                                        const unsigned char* __restrict__ dof, int filter_dbc,
                                                        double* __restrict__ val, double* __restrict__ dinv,
                                                        double* __restrict__ diag) {
  const int j0 = blockIdx.x * kBlock, ub = j0 + threadIdx.x;
  double S[9];
  staged_run_sum<9>(hseg, j0, min(j0 + kBlock, nub), hcontrib, S);
  if (ub >= nub) return;
  hess_store_block(ub, ub_rc[ub], S, ub_pos, dof, filter_dbc, val, dinv, diag);
}

void launch_hess_reduce(int nub, const int* hseg, const double* hcontrib, const int2* ub_rc, const int* ub_pos,
                        const unsigned char* dof, int filter_dbc, double* bell_val, double* dinv, double* diag,
                        hipStream_t s) {
  if (nub <= 0) return;
  k_hess_reduce<<<nblocks(nub), kBlock, 0, s>>>(nub, hseg, hcontrib, ub_rc, ub_pos, dof, filter_dbc, bell_val, dinv,
                                                diag);
}

// Tile-reduced Hessian assembly: workgroup b evaluates the kHessTile tets of tile b and writes one partial per
// (tile, upper block) and component (device_common.h::elem_hess_tile); k_hess_gather adds each block's partials in
// tile order.  Replaces the per-tet slots of k_elem_hess + k_hess_reduce (10 blocks of 72 B per tet written and
// read back) by about 2.8 partial blocks per tet on a structured bar.
template <typename T, int Model>
__global__ __launch_bounds__(kHessTile) void k_elem_hess_tile(MeshView m, const double* __restrict__ u, double dt2,
                                                              int make_spsd, const unsigned* __restrict__ hslot,
                                                              const int* __restrict__ ht_off,
                                                              const int* __restrict__ ht_run,
                                                              double* __restrict__ hpart, long long P) {
  const int nt = m.nt > 0 ? m.nt : 1;
  const int lb = logical_block((nt + kHessTile - 1) / kHessTile, m.xcd_remap);
  if (lb < 0) return;  // uniform over the workgroup
  const int e = lb * kHessTile + threadIdx.x;
  elem_hess_tile<T, Model, kHessTile>(m, lb, e < m.nt && m.cells[e].x >= 0, u, dt2, make_spsd, hslot, ht_off, ht_run,
                                      hpart, P);
}

void launch_elem_hess_tile(const MeshView& m, const double* u, double dt2, int make_spsd, const unsigned* hslot,
                           const int* ht_off, const int* ht_run, double* hpart, long long P, bool fp32,
                           hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  if (m.nt <= 0) return;
  const int nb = (m.nt + kHessTile - 1) / kHessTile;
  dispatch_elem(m.model, fp32, [&](auto k) {
    using K = decltype(k);
    launch_timed(k_elem_hess_tile<typename K::type, K::model>, grid_for(nb, m.xcd_remap), kHessTile, s, e0, e1, m, u,
                 dt2, make_spsd, hslot, ht_off, ht_run, hpart, P);
  });
}

// One thread per upper block (r ≤ c): the sum of its tile partials, then as k_hess_reduce (Dirichlet identity
// filter, blocked-ELL store, Jacobi inverse diagonal).
__global__ __launch_bounds__(kBlock) void k_hess_gather(int nub, const int* __restrict__ hb_seg,
                                                        const int* __restrict__ hb_pos,
                                                        const double* __restrict__ hpart, long long P,
                                                        const int2* __restrict__ ub_rc, const int* __restrict__ ub_pos,
                                                        const unsigned char* __restrict__ dof, int filter_dbc,
                                                        double* __restrict__ val, double* __restrict__ dinv,
                                                        double* __restrict__ diag) {
  const int ub = blockIdx.x * kBlock + threadIdx.x;
  if (ub >= nub) return;
  double S[9];
  hess_gather_block(ub, hb_seg, hb_pos, hpart, P, S);
  hess_store_block(ub, ub_rc[ub], S, ub_pos, dof, filter_dbc, val, dinv, diag);
}

void launch_hess_gather(int nub, const int* hb_seg, const int* hb_pos, const double* hpart, long long P,
                        const int2* ub_rc, const int* ub_pos, const unsigned char* dof, int filter_dbc,
                        double* bell_val, double* dinv, double* diag, hipStream_t s, hipEvent_t e0,
                        hipEvent_t e1) {
  if (nub <= 0) return;
  launch_timed(k_hess_gather, nblocks(nub), kBlock, s, e0, e1, nub, hb_seg, hb_pos, hpart, P, ub_rc, ub_pos, dof,
               filter_dbc, bell_val, dinv, diag);
}

// per-element outputs for parity checks: Ψ, P, H (full 9x9, projected if make_spsd), K (full 12x12, unscaled)
template <typename T, int Model>
__global__ __launch_bounds__(kBlock) void k_elem_eval(MeshView m, const double* __restrict__ u, int make_spsd,
                                                      double* psi, double* Pout, double* Hout, double* Kout) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= m.nt) return;
  const TetLocal t = load_tet(m, e);
  double x[4][3];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const size_t k = 3 * (size_t)t.id[i] + a;
      x[i][a] = m.X[k] + u[k];
    }
  double Fd[9];
  def_grad(x, t.B, Fd);
  T F[9], Bt[9], P[9], H[45], D[4][3];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    F[k] = (T)Fd[k];
    Bt[k] = (T)t.B[k];
  }
  const T ps = material_stress<Model>(F, (T)t.lam, (T)t.mu, P);
  if (psi) psi[e] = (double)ps;
  if (Pout)
    for (int k = 0; k < 9; ++k) Pout[9 * (size_t)e + k] = (double)P[k];
  if (Hout || Kout) {
    material_hessian<Model>(F, (T)t.lam, (T)t.mu, make_spsd, (T)kSpdFloor, H);
    if (Hout)
      for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) Hout[81 * (size_t)e + i * 9 + j] = (double)H[sym9(i, j)];
    if (Kout) {
      shape_grads(Bt, D);
      double* Ke = Kout + 144 * (size_t)e;
      pullback_blocks(H, D, [&](int v, int w, const T K[9]) {
        for (int b = 0; b < 3; ++b)
          for (int dd = 0; dd < 3; ++dd) {
            Ke[(3 * v + b) * 12 + 3 * w + dd] = (double)K[b * 3 + dd];
            Ke[(3 * w + dd) * 12 + 3 * v + b] = (double)K[b * 3 + dd];
          }
      });
    }
  }
}

void launch_elem_eval(const MeshView& m, const double* u, int make_spsd, double* psi, double* P, double* H,
                      double* K, bool fp32, hipStream_t s) {
  if (m.nt <= 0) return;
  dispatch_elem(m.model, fp32, [&](auto k) {
    using Kd = decltype(k);
    k_elem_eval<typename Kd::type, Kd::model><<<nblocks(m.nt), kBlock, 0, s>>>(m, u, make_spsd, psi, P, H, K);
  });
}

// ------------------------------------------------------------------------------------------------------
// batched material evaluation (vmap over the models' stress / Hessian operators), one item per thread

template <int DIM, int Model>
__global__ __launch_bounds__(kBlock) void k_material_eval(long long n, const double* __restrict__ F,
                                                          const double* __restrict__ lam, const double* __restrict__ mu,
                                                          int make_spsd, double* __restrict__ psi,
                                                          double* __restrict__ P, double* __restrict__ H) {
  constexpr int D2 = DIM * DIM;
  __shared__ double lds[64 * (D2 * D2 | 1)];
  const long long first = (long long)blockIdx.x * kBlock;
  const long long i = first + threadIdx.x;
  const int count = n - first < kBlock ? (int)(n - first) : kBlock;
  const bool live = i < n;
  // every lane runs the arithmetic (the staged stores synchronise the workgroup); lanes past the end use F = I
  double f[D2], p[D2];
#pragma unroll
  for (int k = 0; k < D2; ++k) f[k] = live ? F[i * D2 + k] : (k % (DIM + 1) == 0 ? 1.0 : 0.0);
  const double l = live ? lam[i] : 1.0, m = live ? mu[i] : 1.0;
  double e;
  if constexpr (DIM == 2) {
    e = Model == kLinearElastic ? linear2_stress(f, l, m, p) : snh2_stress(f, l, m, p);
    if (H) {
      double h[16];
      if (Model == kLinearElastic) {
        linear2_hessian(l, m, h);
        if (make_spsd) sym4_project_spd(h, kSpdFloor);
      } else if (make_spsd) {
        snh2_project_spd(f, l, m, kSpdFloor, h);
      } else {
        snh2_hessian(f, l, m, h);
      }
      store_items<16>(H, first, count, lds, [&](int k) { return h[k]; });
    }
  } else {
    e = material_stress<Model>(f, l, m, p);
    if (H) {
      double h[45];
      material_hessian<Model>(f, l, m, make_spsd, kSpdFloor, h);
      store_items<81>(H, first, count, lds, [&](int k) { return h[sym9(k / 9, k % 9)]; });
    }
  }
  if (live && psi) psi[i] = e;
  if (P) {
    if constexpr (DIM == 2) {
      if (live)
#pragma unroll
        for (int k = 0; k < D2; ++k) P[i * D2 + k] = p[k];
    } else {
      store_items<9>(P, first, count, lds, [&](int k) { return p[k]; });
    }
  }
}

void launch_material_eval(int dim, int model, long long n, const double* F, const double* lam, const double* mu,
                          int make_spsd, double* psi, double* P, double* H, hipStream_t s) {
  if (n <= 0) return;
  const int nb = (int)((n + kBlock - 1) / kBlock);
  if (dim == 2) {
    if (model == kLinearElastic)
      k_material_eval<2, kLinearElastic><<<nb, kBlock, 0, s>>>(n, F, lam, mu, make_spsd, psi, P, H);
    else
      k_material_eval<2, kStableNeoHookean><<<nb, kBlock, 0, s>>>(n, F, lam, mu, make_spsd, psi, P, H);
  } else {
    if (model == kLinearElastic)
      k_material_eval<3, kLinearElastic><<<nb, kBlock, 0, s>>>(n, F, lam, mu, make_spsd, psi, P, H);
    else
      k_material_eval<3, kStableNeoHookean><<<nb, kBlock, 0, s>>>(n, F, lam, mu, make_spsd, psi, P, H);
  }
}

// ------------------------------------------------------------------------------------------------------
// P1 triangles in 2D: the per-element quantities of basic_time_step<..., PhysicalDim 2, TopologyDim 3>.  One
// thread per triangle.  With D_a the physical gradient of vertex a's hat function (D_0 = −(row 0 + row 1) of
// Dm⁻¹, D_a = row a−1 for a = 1, 2), ∂F_ij/∂x_ac = δ_ic·D_a(j) (compute_pfpx_2, elast/internal_physic.hpp:78-100),
// so the vertex gradient is A·P·D_a, the stiffness block (a c, b d) is A·Σ_jl D_a(j)·H[(c j), (d l)]·D_b(l), the
// consistent mass is ρA/12·(1 + δ_ab) and the Laplacian A·D_a·D_b (bilinears.hpp:65-91 with the tables of
// :27-41).  A triangle with a vertex index outside [0, nv) is not read: its outputs are NaN.

template <int Model>
__global__ __launch_bounds__(kBlock) void k_tri_eval(int nv, int nt, const double* __restrict__ X,
                                                     const int* __restrict__ cells, const double* __restrict__ u,
                                                     const double* __restrict__ lam, const double* __restrict__ mu,
                                                     double rho, int make_spsd, TriOut o) {
  __shared__ double lds[64 * 37];
  const int e = blockIdx.x * kBlock + threadIdx.x;
  const long long first = (long long)blockIdx.x * kBlock;
  const int count = nt - (int)first < kBlock ? nt - (int)first : kBlock;
  const bool live = e < nt;
  int c[3] = {0, 0, 0};
  bool ok = live;
  if (live) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      c[a] = cells[3 * e + a];
      ok = ok && c[a] >= 0 && c[a] < nv;
    }
  }
  // every lane runs the arithmetic (the stores below synchronise the workgroup); a triangle that is not read
  // computes from zeros and reports NaN
  double x[3][2] = {}, y[3][2] = {};
  double l = 1.0, m = 1.0;
  if (ok) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        x[a][i] = X[2 * c[a] + i];
        y[a][i] = x[a][i] + u[2 * c[a] + i];
      }
    l = lam[e];
    m = mu[e];
  }
  const double bad = ok ? 0.0 : __builtin_nan("");  // added to every output: NaN for a triangle that was not read
  // Dm = [x1 − x0, x2 − x0] by columns, its inverse by the adjugate
  const double m00 = x[1][0] - x[0][0], m01 = x[2][0] - x[0][0], m10 = x[1][1] - x[0][1], m11 = x[2][1] - x[0][1];
  const double det = m00 * m11 - m01 * m10;
  const double Di[4] = {m11 / det, -m01 / det, -m10 / det, m00 / det};
  const double A = 0.5 * fabs(det);
  double f[4];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double s0 = y[1][i] - y[0][i], s1 = y[2][i] - y[0][i];
#pragma unroll
    for (int j = 0; j < 2; ++j) f[2 * i + j] = s0 * Di[j] + s1 * Di[2 + j];
  }
  double P[4];
  const double psi = Model == kLinearElastic ? linear2_stress(f, l, m, P) : snh2_stress(f, l, m, P);
  double D[3][2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    D[1][j] = Di[j];
    D[2][j] = Di[2 + j];
    D[0][j] = -(Di[j] + Di[2 + j]);
  }
  if (live) {
    if (o.area) o.area[e] = A + bad;
    if (o.energy) o.energy[e] = A * psi + bad;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (o.dminv) o.dminv[4 * (long long)e + k] = Di[k] + bad;
      if (o.F) o.F[4 * (long long)e + k] = f[k] + bad;
    }
  }
  if (o.grad)
    store_items<6>(o.grad, first, count, lds, [&](int k) {
      const int a = k / 2, i = k % 2;
      return A * (P[2 * i] * D[a][0] + P[2 * i + 1] * D[a][1]) + bad;
    });
  if (o.mass)
    store_items<9>(o.mass, first, count, lds,
                   [&](int k) { return rho * A * (k / 3 == k % 3 ? 1.0 / 6.0 : 1.0 / 12.0) + bad; });
  if (o.laplacian)
    store_items<9>(o.laplacian, first, count, lds, [&](int k) {
      const int a = k / 3, b = k % 3;
      return A * (D[a][0] * D[b][0] + D[a][1] * D[b][1]) + bad;
    });
  if (o.stiff) {
    double H[16];
    if (Model == kLinearElastic) {
      linear2_hessian(l, m, H);
      if (make_spsd) sym4_project_spd(H, kSpdFloor);
    } else if (make_spsd) {
      snh2_project_spd(f, l, m, kSpdFloor, H);
    } else {
      snh2_hessian(f, l, m, H);
    }
    // K[(a c), (b d)] = A·Σ_jl D_a(j)·H[(c j), (d l)]·D_b(l)
    store_items<36>(o.stiff, first, count, lds, [&](int k) {
      const int a = k / 12, ci = (k / 6) % 2, b = (k % 6) / 2, di = k % 2;
      double v = 0;
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int q = 0; q < 2; ++q) v += D[a][j] * H[4 * (2 * ci + j) + 2 * di + q] * D[b][q];
      return A * v + bad;
    });
  }
}

void launch_tri_eval(int model, int nv, int nt, const double* X, const int* cells, const double* u,
                     const double* lam, const double* mu, double rho, int make_spsd, const TriOut& o, hipStream_t s) {
  if (nt <= 0) return;
  if (model == kLinearElastic)
    k_tri_eval<kLinearElastic><<<nblocks(nt), kBlock, 0, s>>>(nv, nt, X, cells, u, lam, mu, rho, make_spsd, o);
  else
    k_tri_eval<kStableNeoHookean><<<nblocks(nt), kBlock, 0, s>>>(nv, nt, X, cells, u, lam, mu, rho, make_spsd, o);
}

// ------------------------------------------------------------------------------------------------------
// adjoint of one step (element part) and the Dirichlet mask of its right-hand side

template <int Model>
__global__ __launch_bounds__(kBlock) void k_elem_adjoint(MeshView m, const double* __restrict__ u,
                                                         const double* __restrict__ lv, double dt2,
                                                         const int* __restrict__ gperm, double* __restrict__ gcontrib,
                                                         double* __restrict__ g_lam, double* __restrict__ g_mu) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e < m.nt) elem_adjoint_tet<Model>(m, e, u, lv, dt2, gperm, gcontrib, g_lam, g_mu);
}

void launch_elem_adjoint(const MeshView& m, const double* u, const double* lv, double dt2, const int* gperm,
                         double* gcontrib, double* g_lam, double* g_mu, hipStream_t s) {
  if (m.nt <= 0) return;
  if (m.model == kLinearElastic)
    k_elem_adjoint<kLinearElastic><<<nblocks(m.nt), kBlock, 0, s>>>(m, u, lv, dt2, gperm, gcontrib, g_lam, g_mu);
  else
    k_elem_adjoint<kStableNeoHookean><<<nblocks(m.nt), kBlock, 0, s>>>(m, u, lv, dt2, gperm, gcontrib, g_lam, g_mu);
}

__global__ void k_mask_dbc(int n, const unsigned char* __restrict__ dof, const double* __restrict__ in,
                           double* __restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = dof[i] ? 0.0 : in[i];
}

void launch_mask_dbc(int n, const unsigned char* dof, const double* in, double* out, hipStream_t s) {
  if (n <= 0) return;
  k_mask_dbc<<<nblocks(n), kBlock, 0, s>>>(n, dof, in, out);
}

// ------------------------------------------------------------------------------------------------------
// Symmetric blocked-ELL SpMV.  Each block row is handled by LPR lanes (LPR = 1 or 4; lane j takes slots j,
// j+LPR, ...).  Padded slots hold a valid column (the row itself) and zero values, and padded lower references
// point at a zero block, so the slot loops are branch-free with a trip count uniform over the slice and can be
// software-pipelined.  With LPR = 4 a wave covers 16 rows: lane l takes row l & 15 and slot phase j = l >> 4, so
// every 16-lane group reads 16 consecutive doubles; the 4 partial results of a row are combined with two
// shuffles in a fixed order.  Upper blocks are read from the row's own slots, lower blocks as transposes of the
// stored (w, v) blocks.

template <int LPR>
__global__ __launch_bounds__(kSpmvBlock) void k_spmv(BellView A, const double* __restrict__ x, double* __restrict__ y) {
  const int lb = logical_block(RowMap<LPR>::blocks(A.nrows), A.xcd_remap);
  if (lb < 0) return;
  int v, j;
  RowMap<LPR>::map(lb, v, j);
  double r[3] = {0.0, 0.0, 0.0};
  if (v < A.nrows) row_part<LPR>(A, v, j, x, r);
  row_combine<LPR>(r);
  if (v < A.nrows) {
    if (LPR <= 1) {
      y[3 * (size_t)v] = r[0];
      y[3 * (size_t)v + 1] = r[1];
      y[3 * (size_t)v + 2] = r[2];
    } else if (j < 3) {
      y[3 * (size_t)v + j] = r[j];
    }
  }
}

void launch_spmv(const BellView& A, const double* x, double* y, hipStream_t s) {
  if (A.nrows <= 0) return;
  if (A.lanes_per_row == 4)
    k_spmv<4><<<grid_for(RowMap<4>::blocks(A.nrows), A.xcd_remap), kSpmvBlock, 0, s>>>(A, x, y);
  else if (A.lanes_per_row == 0)
    k_spmv<0><<<grid_for(RowMap<0>::blocks(A.nrows), A.xcd_remap), kSpmvBlock, 0, s>>>(A, x, y);
  else
    k_spmv<1><<<grid_for(RowMap<1>::blocks(A.nrows), A.xcd_remap), kSpmvBlock, 0, s>>>(A, x, y);
}

constexpr int kVec = 4;  // elements per thread of the vector kernels
static inline int vblocks(int n) { return (int)(((long long)(n > 0 ? n : 1) + kBlock * kVec - 1) / (kBlock * kVec)); }

__global__ __launch_bounds__(kBlock) void k_dot(int n, const double* __restrict__ a, const double* __restrict__ b,
                                                double* __restrict__ part) {
  double v[1] = {0.0};
#pragma unroll
  for (int j = 0; j < kVec; ++j) {
    const int i = (blockIdx.x * kVec + j) * kBlock + threadIdx.x;
    if (i < n) v[0] += a[i] * b[i];
  }
  store_partials<1>(v, part);
}

void launch_dot(int n, const double* a, const double* b, double* part, double* out, hipStream_t s) {
  const int nb = vblocks(n);
  k_dot<<<nb, kBlock, 0, s>>>(n, a, b, part);
  launch_finalize(part, nb, 1, nullptr, 0, out, s);
}

// ------------------------------------------------------------------------------------------------------
// vector kernels of the step (time_step.hpp:499-569)

__global__ void k_prepare(int n, const double* __restrict__ u, const double* __restrict__ vel,
                          const double* __restrict__ accel, const unsigned char* __restrict__ dof,
                          const double* __restrict__ dbc, double dt, double* __restrict__ uhat,
                          double* __restrict__ unext) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double uh = u[i];
  uh += dt * vel[i];
  uh += dt * dt * accel[i];
  if (dof[i]) uh = dbc[i];
  uhat[i] = uh;
  unext[i] = uh;
}

void launch_prepare(int n, const double* u, const double* vel, const double* accel, const unsigned char* dof,
                    const double* dbc, double dt, double* uhat, double* unext, hipStream_t s) {
  if (n <= 0) return;
  k_prepare<<<nblocks(n), kBlock, 0, s>>>(n, u, vel, accel, dof, dbc, dt, uhat, unext);
}

__global__ void k_step_next(int n, const double* __restrict__ unext, double* __restrict__ u,
                            double* __restrict__ uprev, double* __restrict__ vel, double dt, double damping) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double cur = u[i], nxt = unext[i];
  uprev[i] = cur;
  u[i] = nxt;
  double v = nxt;
  v += -1.0 * cur;
  v *= 1.0 / dt;
  if (damping != 0.0) v *= (1.0 - damping);
  vel[i] = v;
}

void launch_step_next(int n, const double* unext, double* u, double* uprev, double* vel, double dt, double damping,
                      hipStream_t s) {
  if (n <= 0) return;
  k_step_next<<<nblocks(n), kBlock, 0, s>>>(n, unext, u, uprev, vel, dt, damping);
}

__global__ void k_restore(int n, const unsigned char* __restrict__ dof, const double* __restrict__ dbc,
                          double* __restrict__ u, double* __restrict__ uprev, double* __restrict__ vel, double dt) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double up = uprev[i], cu = u[i];
  if (dof[i]) {
    up = dbc[i];
    cu = dbc[i];
  }
  uprev[i] = up;
  u[i] = cu;
  double v = cu;
  v += -1.0 * up;
  v *= 1.0 / dt;
  vel[i] = v;
}

void launch_restore(int n, const unsigned char* dof, const double* dbc, double* u, double* uprev, double* vel,
                    double dt, hipStream_t s) {
  if (n <= 0) return;
  k_restore<<<nblocks(n), kBlock, 0, s>>>(n, dof, dbc, u, uprev, vel, dt);
}

__global__ void k_axpy_to(int n, const double* __restrict__ x, double alpha, const double* __restrict__ d,
                          double* __restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = x[i] + alpha * d[i];
}

void launch_axpy_to(int n, const double* x, double alpha, const double* d, double* out, hipStream_t s) {
  if (n <= 0) return;
  k_axpy_to<<<nblocks(n), kBlock, 0, s>>>(n, x, alpha, d, out);
}

// ------------------------------------------------------------------------------------------------------
// Jacobi-PCG (Hestenes–Stiefel), x0 = 0, stop when ‖r‖₂² ≤ rtol²·‖b‖₂² or after maxit iterations, in two
// kernels per iteration.  Scalars never leave the device: the (ρ = r·z, r·r) partials of residual k live in
// pb[k%2] (the host alternates the buffers), p alternates between two vectors, and once the system is solved
// every later kernel of the chunk returns at entry.

__global__ __launch_bounds__(kBlock) void k_pcg_init(int n, const double* __restrict__ b, double sign,
                                                     const double* __restrict__ dinv, double* __restrict__ x,
                                                     double* __restrict__ r, double* __restrict__ z,
                                                     double* __restrict__ pb) {
  double v[2] = {0.0, 0.0};
#pragma unroll
  for (int j = 0; j < kVec; ++j) {
    const int i = (blockIdx.x * kVec + j) * kBlock + threadIdx.x;
    if (i < n) {
      const double bi = sign * b[i];
      const double zi = dinv[i] * bi;
      x[i] = 0.0;
      r[i] = bi;
      z[i] = zi;
      v[0] += bi * zi;
      v[1] += bi * bi;
    }
  }
  store_partials<2>(v, pb);
}

__global__ __launch_bounds__(kBlock) void k_pcg_init_fin(const double* __restrict__ pb, int nb, double rtol,
                                                         int maxit, PcgScalars* sc, PcgIter* it) {
  double t[2];
  sum_partials<2>(pb, nb, t);
  if (threadIdx.x == 0) {
    it->ka = 0;
    it->kb = 0;
    sc->rz = t[0];
    sc->rz_hist[0] = t[0];
    sc->bb = t[1];
    sc->rr = t[1];
    sc->tol2 = rtol * rtol * t[1];
    sc->iters = 0;
    sc->maxit = maxit;
    sc->breakdown = 0;
    sc->done = (t[1] <= sc->tol2 || maxit <= 0 || !(t[1] == t[1])) ? 1 : 0;
    sc->stop_k = sc->done ? -1 : INT_MAX;
  }
}

int launch_pcg_init(int n, const double* b, double sign, const double* dinv, double* x, double* r, double* z,
                    double rtol, int maxit, PcgScalars* sc, PcgIter* it, double* pb, hipStream_t s) {
  const int nb = vblocks(n);
  k_pcg_init<<<nb, kBlock, 0, s>>>(n, b, sign, dinv, x, r, z, pb);
  k_pcg_init_fin<<<1, kBlock, 0, s>>>(pb, nb, rtol, maxit, sc, it);
  return nb;
}

// Iteration k, kernel A: s = A z_k, then (after the loop, so the partial-sum latency overlaps other
// workgroups' SpMV work) the convergence test on ‖r_k‖² from the previous update's partials, x_k = x_{k−1} +
// α_{k−1}·p_{k−1} (α published by the previous kernel B; applied here, where p_{k−1} is read anyway, also on the
// iteration that stops), β = ρ_k/ρ_{k−1}, p_k = z_k + β p_{k−1} and q_k = A p_k = s + β q_{k−1} for the lane's
// own row (in place), and the partials of p_k·q_k.  For k = 0: p_0 = z_0, q_0 = s.  k is read from the device counter (kernel A reads ka and publishes
// kb = ka, kernel B reads kb and publishes ka = kb + 1), so every iteration's launch is identical and a chunk of
// iterations can be replayed from a captured graph; the (r·z, r·r) partials of residual k live in pb[k % 2], and
// ρ_{k−1} is read from the slot the previous launch wrote (the same sum, so only one partial array is summed).
// At 786k tets the SpMV has about 2.2 waves per SIMD: a register budget for 3 waves keeps the whole grid in
// one dispatch round while leaving the compiler room to keep many of a row's loads in flight.
template <int LPR>
__global__ __launch_bounds__(kSpmvBlock) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_pcg_spmv(
    BellView A, const double* __restrict__ z, double* __restrict__ p, double* __restrict__ q,
    const double* __restrict__ pb0, const double* __restrict__ pb1, int nb, PcgIter* __restrict__ it,
    PcgScalars* __restrict__ sc, double* __restrict__ pa, double* __restrict__ x) {
  // The iteration counter and every scalar this launch reads, in one round of loads (none depends on another;
  // ρ_{k−1} is selected from both slots afterwards).  The empty asm statement takes them as inputs, which keeps
  // the compiler from sinking the loads below the row product.
  const int lb = logical_block(RowMap<LPR>::blocks(A.nrows), A.xcd_remap);
  int v = 0, j = 0;
  if (lb >= 0) RowMap<LPR>::map(lb, v, j);
  const bool active = lb >= 0 && v < A.nrows;
  // the fixed layout's index words do not depend on the scalars: their loads go out with them
  FixedRowIndex ix{};
  if (LPR == 0 && active) ix = bell_fixed_index(A, v);
  const int k = it->ka;
  const int stop_k = sc->stop_k, maxit = sc->maxit;
  const double tol2 = sc->tol2, al = sc->alpha;  // α_{k−1} (used only for k > 0)
  const double rz0 = sc->rz_hist[0], rz1 = sc->rz_hist[1];
  asm volatile("" ::"s"(k), "s"(stop_k), "s"(maxit), "s"(tol2), "s"(al), "s"(rz0), "s"(rz1));
  const double rz_prev = (k & 1) ? rz0 : rz1;  // ρ_{k−1} (slot (k − 1) % 2), published by the previous launch
  if (blockIdx.x == 0 && threadIdx.x == 0) it->kb = k;
  // Only a stop from an earlier launch ends the work here: workgroup 0 of this launch may publish a stop at
  // iteration k while other workgroups have yet to start, and those must still apply x_k = x_{k−1} + α p for
  // their rows (below) -- they reach the same stop decision from the same partial sums.
  if (k > stop_k) return;
  const double* __restrict__ pb_cur = (k & 1) ? pb1 : pb0;
  double sv[3] = {0.0, 0.0, 0.0};
  if (active) {
    if (LPR == 0)
      bell_row_fixed(A, v, ix, z, sv);
    else
      row_part<LPR>(A, v, j, z, sv);
  }
  row_combine<LPR>(sv);
  // The lane's own entries of z, p_{k−1}, q_{k−1} and x and the partial sums of residual k are loaded in one
  // round, so the rest of the kernel waits for one memory latency instead of one per phase.
  double zo[3] = {0.0, 0.0, 0.0}, po[3] = {0.0, 0.0, 0.0}, qo[3] = {0.0, 0.0, 0.0}, xo[3] = {0.0, 0.0, 0.0};
  if (active) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (LPR <= 1 || a == j) {  // with LPR = 4, lane j of the row owns component j
        const size_t i = 3 * (size_t)v + a;
        zo[a] = z[i];
        po[a] = p[i];
        qo[a] = q[i];
        xo[a] = x[i];
      }
  }
  double t_cur[2] = {0.0, 0.0};
  if (k > 0) sum_partials<2, kSpmvBlock>(pb_cur, nb, t_cur);
  double beta = 0.0;
  if (k > 0) {
    const bool nan = !(t_cur[1] == t_cur[1]);
    const bool done = t_cur[1] <= tol2 || k >= maxit || nan;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->iters = k;
      sc->rr = t_cur[1];
      sc->rz = t_cur[0];
      sc->rz_hist[k & 1] = t_cur[0];
      if (nan) sc->breakdown = 2;
      if (done) {
        sc->stop_k = k;  // launches of later iterations return at entry
        sc->done = 1;    // read by kernel B of this iteration (the next launch) and by the host
      }
    }
    // x_k = x_{k−1} + α_{k−1}·p_{k−1} for the own row (kernel B published α_{k−1})
    if (active) {
#pragma unroll
      for (int a = 0; a < 3; ++a)
        if (LPR <= 1 || a == j) x[3 * (size_t)v + a] = xo[a] + al * po[a];
    }
    if (done) return;
    beta = t_cur[0] / rz_prev;
  }
  double acc[1] = {0.0};
  if (active) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (LPR <= 1 || a == j) {
        const size_t i = 3 * (size_t)v + a;
        double pv = zo[a], qv = sv[a];
        if (k > 0) {
          pv += beta * po[a];
          qv += beta * qo[a];
        }
        p[i] = pv;
        q[i] = qv;
        acc[0] += pv * qv;
      }
    }
  }
  store_partials<1, kSpmvBlock>(acc, pa);
}

// Iteration k, kernel B: α = ρ_k/(p_k·q) with ρ_k from the slot kernel A wrote; r −= α q, z = D⁻¹ r; partials
// of (r·z, r·r); α is published for the next kernel A, which applies x += α p_k.
// The streaming loads are issued before the partial sums so that their latency overlaps.  r and z leave with
// non-temporal stores: no kernel reads them back through this XCD's L2 before the next kernel boundary, and the
// update then ends sooner (6.25 → 5.9 µs per launch, −1.3 % per PCG iteration at 786k tets).
__global__ __launch_bounds__(kBlock) void k_pcg_update(int n, const double* __restrict__ dinv,
                                                       double* __restrict__ r, double* __restrict__ z,
                                                       const double* __restrict__ q,
                                                       const double* __restrict__ pa, int na, double* __restrict__ pb0,
                                                       double* __restrict__ pb1, int nb,
                                                       PcgIter* __restrict__ it, PcgScalars* __restrict__ sc) {
  // The streaming loads go out first, with the counter, done and both ρ slots (ρ_k is the slot kernel A wrote):
  // the kernel then waits for one memory latency before its partial sums instead of two.  A launch after the
  // solve stopped reads its share of r, q and D⁻¹ for nothing (the host's polling bounds those launches).
  double rv[kVec], qv[kVec], dv[kVec];
#pragma unroll
  for (int j = 0; j < kVec; ++j) {  // clamped indices: no branch between the loads (the tail's are not used)
    const int i0 = (blockIdx.x * kVec + j) * kBlock + threadIdx.x;
    const int i = i0 < n ? i0 : n - 1;
    rv[j] = r[i];
    qv[j] = q[i];
    dv[j] = dinv[i];
  }
  const int k = it->kb;
  const int done = sc->done;
  const double rz0 = sc->rz_hist[0], rz1 = sc->rz_hist[1];
  asm volatile("" ::"s"(k), "s"(done), "s"(rz0), "s"(rz1));
  const double rz_k = (k & 1) ? rz1 : rz0;
  if (blockIdx.x == 0 && threadIdx.x == 0) it->ka = k + 1;
  if (done) return;
  const double* __restrict__ pb_cur = (k & 1) ? pb1 : pb0;
  double* __restrict__ pb_next = (k & 1) ? pb0 : pb1;
  double pq[1];
  if (na <= 16 * kBlock)
    sum_partials_1pass<1, kBlock, 16>(pa, na, pq);  // up to 4,096 partials (every mesh up to ~260k vertices)
  else
    sum_partials<1, kBlock, 16>(pa, na, pq);
  if (!(pq[0] > 0.0)) {  // not positive definite (or NaN): every workgroup sees the same pq
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->breakdown = 1;
      sc->stop_k = k;
      sc->done = 1;
    }
    return;
  }
  const double alpha = rz_k / pq[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->alpha = alpha;
  double v[2] = {0.0, 0.0};
#pragma unroll
  for (int j = 0; j < kVec; ++j) {
    const int i = (blockIdx.x * kVec + j) * kBlock + threadIdx.x;
    if (i < n) {
      const double ri = rv[j] - alpha * qv[j];
      const double zi = dv[j] * ri;
      __builtin_nontemporal_store(ri, &r[i]);
      __builtin_nontemporal_store(zi, &z[i]);
      v[0] += ri * zi;
      v[1] += ri * ri;
    }
  }
  store_partials<2>(v, pb_next);
}

int launch_pcg_spmv(const BellView& A, const double* z, double* p, double* q, const double* pb0, const double* pb1,
                    int nb, PcgIter* it, PcgScalars* sc, double* pa, double* x, hipStream_t s, hipEvent_t e0,
                    hipEvent_t e1) {
  const int rows = A.nrows > 0 ? A.nrows : 1;
  if (A.lanes_per_row == 4) {
    const int grid = grid_for(RowMap<4>::blocks(rows), A.xcd_remap);
    launch_timed(k_pcg_spmv<4>, grid, kSpmvBlock, s, e0, e1, A, z, p, q, pb0, pb1, nb, it, sc, pa, x);
    return grid;
  }
  if (A.lanes_per_row == 0) {
    const int grid = grid_for(RowMap<0>::blocks(rows), A.xcd_remap);
    launch_timed(k_pcg_spmv<0>, grid, kSpmvBlock, s, e0, e1, A, z, p, q, pb0, pb1, nb, it, sc, pa, x);
    return grid;
  }
  const int grid = grid_for(RowMap<1>::blocks(rows), A.xcd_remap);
  launch_timed(k_pcg_spmv<1>, grid, kSpmvBlock, s, e0, e1, A, z, p, q, pb0, pb1, nb, it, sc, pa, x);
  return grid;
}

void launch_pcg_update(int n, const double* dinv, double* r, double* z, const double* q,
                       const double* pa, int na, double* pb0, double* pb1, PcgIter* it, PcgScalars* sc,
                       hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  const int nb = vblocks(n);
  launch_timed(k_pcg_update, nb, kBlock, s, e0, e1, n, dinv, r, z, q, pa, na, pb0, pb1, nb, it, sc);
}

int pcg_vector_blocks(int n) { return vblocks(n); }

int pcg_spmv_grid(const BellView& A) {
  const int rows = A.nrows > 0 ? A.nrows : 1;
  if (A.lanes_per_row == 4) return grid_for(RowMap<4>::blocks(rows), A.xcd_remap);
  if (A.lanes_per_row == 0) return grid_for(RowMap<0>::blocks(rows), A.xcd_remap);
  return grid_for(RowMap<1>::blocks(rows), A.xcd_remap);
}

// ------------------------------------------------------------------------------------------------------
// nonlinear-CG helpers

// partials of (dᵀy, gtᵀy, yᵀy, gtᵀd, dᵀd) with y = gt − g
// With a preconditioner P = L·Lᵀ, a = Lᵀ·g_new and b = Lᵀ·g give g_newᵀ·P·y = aᵀ(a − b) and yᵀ·P·y = ‖a − b‖²;
// without one the caller passes a = g_new, b = g.
__global__ __launch_bounds__(kBlock) void k_ncg_dots(int n, const double* __restrict__ gt, const double* __restrict__ g,
                                                     const double* __restrict__ d, const double* __restrict__ a,
                                                     const double* __restrict__ b, double* __restrict__ part) {
  double v[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < kVec; ++j) {
    const int i = (blockIdx.x * kVec + j) * kBlock + threadIdx.x;
    if (i < n) {
      const double gi = gt[i], di = d[i], y = gi - g[i];
      const double ai = a[i], ya = ai - b[i];
      v[0] += di * y;
      v[1] += ai * ya;
      v[2] += ya * ya;
      v[3] += gi * di;
      v[4] += di * di;
    }
  }
  store_partials<5>(v, part);
}

__global__ __launch_bounds__(kBlock) void k_finalize5(const double* __restrict__ part, int np, double* __restrict__ out) {
  double t[5];
  sum_partials<5>(part, np, t);
  if (threadIdx.x < 5) out[threadIdx.x] = t[threadIdx.x];
}

void launch_ncg_dots(int n, const double* gt, const double* g, const double* d, const double* a, const double* b,
                     double* part, double* out, hipStream_t s) {
  const int nb = vblocks(n);
  k_ncg_dots<<<nb, kBlock, 0, s>>>(n, gt, g, d, a, b, part);
  k_finalize5<<<1, kBlock, 0, s>>>(part, nb, out);
}

// Block-CSR product with a general matrix on the block adjacency (blocks row-major, 9 doubles each):
// y_v = Σ_k B_k·x_{w_k} over the neighbours w_k = bcol[k] of v, with B_k = block k (TRANS = false) or the
// transpose of block tpos[k] = (w_k, v) (TRANS = true, i.e. y = Lᵀx).  One thread per block row; used by the
// externally preconditioned NCG (z = L·Lᵀ·g).
template <bool TRANS>
__global__ __launch_bounds__(kBlock) void k_bsr_spmv(int nv, const int* __restrict__ brow, const int* __restrict__ bcol,
                                                     const int* __restrict__ tpos, const double* __restrict__ val,
                                                     const double* __restrict__ x, double* __restrict__ y) {
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= nv) return;
  double y0 = 0, y1 = 0, y2 = 0;
  for (int k = brow[v]; k < brow[v + 1]; ++k) {
    const size_t w3 = 3 * (size_t)bcol[k];
    const double x0 = x[w3], x1 = x[w3 + 1], x2 = x[w3 + 2];
    const double* B = val + 9 * (size_t)(TRANS ? tpos[k] : k);
    if (TRANS) {
      y0 += B[0] * x0 + B[3] * x1 + B[6] * x2;
      y1 += B[1] * x0 + B[4] * x1 + B[7] * x2;
      y2 += B[2] * x0 + B[5] * x1 + B[8] * x2;
    } else {
      y0 += B[0] * x0 + B[1] * x1 + B[2] * x2;
      y1 += B[3] * x0 + B[4] * x1 + B[5] * x2;
      y2 += B[6] * x0 + B[7] * x1 + B[8] * x2;
    }
  }
  y[3 * (size_t)v] = y0;
  y[3 * (size_t)v + 1] = y1;
  y[3 * (size_t)v + 2] = y2;
}

void launch_bsr_spmv(int nv, const int* brow, const int* bcol, const int* tpos, const double* val, const double* x,
                     double* y, bool trans, hipStream_t s) {
  if (nv <= 0) return;
  if (trans)
    k_bsr_spmv<true><<<nblocks(nv), kBlock, 0, s>>>(nv, brow, bcol, tpos, val, x, y);
  else
    k_bsr_spmv<false><<<nblocks(nv), kBlock, 0, s>>>(nv, brow, bcol, tpos, val, x, y);
}

// y = a·x + b·y (b = 0 does not read y)
__global__ void k_axpby(int n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) y[i] = b == 0.0 ? a * x[i] : a * x[i] + b * y[i];
}

void launch_axpby(int n, double a, const double* x, double b, double* y, hipStream_t s) {
  if (n <= 0) return;
  k_axpby<<<nblocks(n), kBlock, 0, s>>>(n, a, x, b, y);
}

// ------------------------------------------------------------------------------------------------------
// L-BFGS in Gram form (H0 = D = diag(h)⁻¹, history of at most kLbfgsSlots pairs in slots 0..m−1)

// Partials of the 6m + 1 dot products one iteration needs (m = kLbfgsSlots, every slot, used or not):
//   [0, m)   s_jᵀg        [m, 2m)  y_jᵀDg       [2m]  gᵀDg
//   [2m+1+j] s_jᵀy_k      [3m+1+j] s_kᵀy_j      [4m+1+j] y_kᵀDy_j      (k = the newest slot; zero when k < 0)
// The newest pair is formed here, s_k = u − u_old and y_k = g − g_old (the accepted iterate and the one before it),
// and stored to slot k by the wave role that reads that slot's other vectors (role 0 stores S_k, role 1 Y_k); the
// other slots are read.  The pair's values are those a separate pass would store, so the sums are unchanged.
// Two wave roles split the dot products by the stored vectors they read, so that each S_j and Y_j is loaded by
// one wave only: role 0 reads S_j, g and y_k (s_jᵀg, s_jᵀy_k: 2m accumulators), role 1 reads Y_j, g, H0, s_k and
// y_k (y_jᵀDg, gᵀDg, s_kᵀy_j, y_kᵀDy_j: 3m + 1 accumulators).  Waves 0 and 1 of a workgroup take one set of
// elements, waves 2 and 3 the next, in a grid-stride loop over a grid of at most kGramMaxBlocks workgroups (all
// resident at once); each trip takes two elements per lane, whose loads are all in flight together.  Partial
// row 2·blockIdx + set holds the workgroup's sums over its element set.
constexpr int kGramMaxBlocks = 768;  // 3 workgroups per CU at 130 VGPRs: all resident
static int gram_grid(int n) { return std::max(1, std::min(kGramMaxBlocks, (n + 255) / 256)); }

__global__ __launch_bounds__(kBlock) void k_lbfgs_gram(int n, double* __restrict__ S, double* __restrict__ Y,
                                                       const double* __restrict__ g, const double* __restrict__ h,
                                                       int k, const double* __restrict__ u,
                                                       const double* __restrict__ u_old,
                                                       const double* __restrict__ g_old, double* __restrict__ part) {
  constexpr int M = kLbfgsSlots;
  static_assert(kBlock == 256, "two element sets of two waves per workgroup");
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int role = wave & 1, set = wave >> 1;
  const int stride = gridDim.x * 128;
  double acc[3 * M];
#pragma unroll
  for (int j = 0; j < 3 * M; ++j) acc[j] = 0.0;
  double gdg = 0.0;
  const bool pair = k >= 0;
  auto body = [&](int i) {
    const double gi = g[i];
    const double sk = pair ? u[i] - u_old[i] : 0.0, yk = pair ? gi - g_old[i] : 0.0;
    if (role == 0) {
      if (pair) S[(size_t)k * n + i] = sk;
#pragma unroll
      for (int j = 0; j < M; ++j) {
        const double sj = pair && j == k ? sk : S[(size_t)j * n + i];
        acc[j] += sj * gi;
        acc[M + j] += sj * yk;
      }
    } else {
      if (pair) Y[(size_t)k * n + i] = yk;
      const double di = 1.0 / h[i], dg = gi * di;
      const double dyk = yk * di;
#pragma unroll
      for (int j = 0; j < M; ++j) {
        const double yj = pair && j == k ? yk : Y[(size_t)j * n + i];
        acc[j] += yj * dg;
        acc[M + j] += sk * yj;
        acc[2 * M + j] += dyk * yj;
      }
      gdg += gi * dg;
    }
  };
  int i = (blockIdx.x * 2 + set) * 64 + lane;
  for (; i + stride < n; i += 2 * stride) {
    body(i);
    body(i + stride);
  }
  if (i < n) body(i);
  double* out = part + (size_t)(blockIdx.x * 2 + set) * kLbfgsDots;
  // role 0: s_jᵀg → [j], s_jᵀy_k → [2m+1+j]; role 1: y_jᵀDg → [m+j], s_kᵀy_j → [3m+1+j], y_kᵀDy_j → [4m+1+j]
#pragma unroll
  for (int j = 0; j < 3 * M; ++j) {
    const double t = wave_sum(acc[j]);
    if (lane != 0) continue;
    const int q = j / M, r = j % M;
    if (role == 0) {
      if (q == 0) out[r] = t;
      else if (q == 1) out[2 * M + 1 + r] = t;
    } else {
      out[(q == 0 ? M : q == 1 ? 3 * M + 1 : 4 * M + 1) + r] = t;
    }
  }
  if (role == 1) {
    const double t = wave_sum(gdg);
    if (lane == 0) out[2 * M] = t;
  }
}

// one workgroup per dot product: the sum of its partials in a fixed order
__global__ __launch_bounds__(kBlock) void k_lbfgs_gram_fin(const double* __restrict__ part, int np,
                                                           double* __restrict__ out) {
  const int d = blockIdx.x;
  double t[1] = {0.0};
  for (int b = threadIdx.x; b < np; b += kBlock) t[0] += part[(size_t)b * kLbfgsDots + d];
  block_sum<1>(t);
  if (threadIdx.x == 0) out[d] = t[0];
}

void launch_lbfgs_gram(int n, double* S, double* Y, const double* g, const double* h, int k, const double* u,
                       const double* u_old, const double* g_old, double* part, double* out, hipStream_t s) {
  k_lbfgs_gram<<<gram_grid(n), kBlock, 0, s>>>(n, S, Y, g, h, k, u, u_old, g_old, part);
  k_lbfgs_gram_fin<<<kLbfgsDots, kBlock, 0, s>>>(part, lbfgs_gram_blocks(n), out);
}

// partial rows written by k_lbfgs_gram: two per workgroup
int lbfgs_gram_blocks(int n) { return 2 * gram_grid(n); }

// new pair and iterate in one pass: s = u_t − u, y = g_t − g, then u = u_t, g = g_t
__global__ __launch_bounds__(kBlock) void k_lbfgs_shift(int n, const double* __restrict__ ut,
                                                        const double* __restrict__ gt, double* __restrict__ u,
                                                        double* __restrict__ g, double* __restrict__ sv,
                                                        double* __restrict__ yv) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double a = ut[i], b = gt[i];
  sv[i] = a - u[i];
  yv[i] = b - g[i];
  u[i] = a;
  g[i] = b;
}

void launch_lbfgs_shift(int n, const double* ut, const double* gt, double* u, double* g, double* sv, double* yv,
                        hipStream_t s) {
  if (n <= 0) return;
  k_lbfgs_shift<<<nblocks(n), kBlock, 0, s>>>(n, ut, gt, u, g, sv, yv);
}

// r = D·(g − Σ_j a_j y_j) + Σ_j c_j s_j over the slots in `mask`, and the line search's first trial point
// u_t = u + α₀·r (the expression of k_axpy_to, so the same bits as that pass would store)
__global__ __launch_bounds__(kBlock) void k_lbfgs_dir(int n, const double* __restrict__ S, const double* __restrict__ Y,
                                                      const double* __restrict__ g, const double* __restrict__ h,
                                                      LbfgsCoef cf, double* __restrict__ r,
                                                      const double* __restrict__ u, double alpha0,
                                                      double* __restrict__ ut) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double q = g[i], t = 0.0;
#pragma unroll
  for (int j = 0; j < kLbfgsSlots; ++j)
    if (cf.mask & (1u << j)) {
      q -= cf.a[j] * Y[(size_t)j * n + i];
      t += cf.c[j] * S[(size_t)j * n + i];
    }
  const double ri = q / h[i] + t;
  r[i] = ri;
  ut[i] = u[i] + alpha0 * ri;
}

void launch_lbfgs_dir(int n, const double* S, const double* Y, const double* g, const double* h, const LbfgsCoef& cf,
                      double* r, const double* u, double alpha0, double* ut, hipStream_t s) {
  if (n <= 0) return;
  k_lbfgs_dir<<<nblocks(n), kBlock, 0, s>>>(n, S, Y, g, h, cf, r, u, alpha0, ut);
}

// ------------------------------------------------------------------------------------------------------
// L-BFGS helpers

__global__ void k_scale_div(int n, const double* __restrict__ q, const double* __restrict__ diag,
                            double* __restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = q[i] / diag[i];
}

void launch_scale_div(int n, const double* q, const double* diag, double* out, hipStream_t s) {
  if (n <= 0) return;
  k_scale_div<<<nblocks(n), kBlock, 0, s>>>(n, q, diag, out);
}

// y += (c1·(*p1) − c2·(*p2))·x with the coefficients read from device memory (p2 may be null)
__global__ void k_axpy_coef(int n, double c1, const double* __restrict__ p1, double c2, const double* __restrict__ p2,
                            const double* __restrict__ x, double* __restrict__ y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const double a = c1 * (*p1);
  const double coef = p2 ? a - c2 * (*p2) : a;
  if (i < n) y[i] += x[i] * coef;
}

void launch_axpy_coef(int n, double c1, const double* p1, double c2, const double* p2, const double* x, double* y,
                      hipStream_t s) {
  if (n <= 0) return;
  k_axpy_coef<<<nblocks(n), kBlock, 0, s>>>(n, c1, p1, c2, p2, x, y);
}

__global__ void k_sub(int n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = a[i] - b[i];
}

void launch_sub(int n, const double* a, const double* b, double* out, hipStream_t s) {
  if (n <= 0) return;
  k_sub<<<nblocks(n), kBlock, 0, s>>>(n, a, b, out);
}

}  // namespace ssim_hip
