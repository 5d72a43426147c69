// device_common.h -- device-side building blocks shared by the kernels of libssim_hip.so: deterministic
// reductions, per-tet gathers and element bodies, and the blocked-ELL row product.
#pragma once
#include <hip/hip_runtime.h>

#include "elem_math.h"
#include "internal.h"

namespace ssim_hip {

// ------------------------------------------------------------------------------------------------------
// reductions

// v of lane (lane pattern Ctrl) in the enabled rows of 16 lanes (RowMask), 0 elsewhere: DPP moves of the two
// 32-bit halves.  Every lane of the wave must be active.
template <int Ctrl, int RowMask = 0xf>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, Ctrl, RowMask, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), Ctrl, RowMask, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over the 64 lanes, returned to every lane, in a fixed order: lane pairs, quads, half rows and rows by DPP
// lane permutations (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror, row_mirror), then rows 0+1 and 2+3
// (row_bcast:15 into rows 1 and 3) and the two halves (row_bcast:31 into row 3); lane 63 holds the total, which
// a read of that lane broadcasts.  DPP moves take a few cycles where a shuffle through ds_bpermute waits for
// the LDS pipeline, and the PCG kernels sit on several of these reductions per iteration.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_move<0xB1>(v);       // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);       // quad_perm [2,3,0,1]
  v += dpp_move<0x141>(v);      // row_half_mirror
  v += dpp_move<0x140>(v);      // row_mirror
  v += dpp_move<0x142, 0xa>(v);  // row_bcast:15
  v += dpp_move<0x143, 0xc>(v);  // row_bcast:31
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Reductions never synchronise workgroups inside a kernel.  A producing kernel stores one partial per
// workgroup (plain stores); the consuming kernel -- the next launch on the stream, so the kernel boundary makes
// the partials visible -- sums all partials in a fixed order in every workgroup.  Every workgroup therefore
// sees the same, bitwise reproducible total.

// block-wide sum of K values over NT threads; the result is returned to every thread
template <int K, int NT = kBlock>
__device__ __forceinline__ void block_sum(double (&v)[K]) {
  static_assert(NT % 64 == 0, "workgroups are whole waves");
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);  // in every lane
  if (NT == 64) return;
  __shared__ double sm[K][NT / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) sm[k][wid] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) s += sm[k][w];
    v[k] = s;
  }
  __syncthreads();
}

// stores this workgroup's partial sums: part[slot*K + k], slot = blockIdx.x unless given
template <int K, int NT = kBlock>
__device__ __forceinline__ void store_partials(double (&v)[K], double* part, int slot = -1) {
  block_sum<K, NT>(v);
  if (slot < 0) slot = blockIdx.x;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) part[(size_t)slot * K + k] = v[k];
  }
}

// total of `count` partials of width K (written by an earlier kernel), in a fixed order.  Each thread loads its
// partials U at a time (independent loads in flight together) before adding them in index order; U = 1 suits
// short runs (at most NT partials), where a deeper buffer only costs registers.
template <int K, int NT = kBlock, int U = 8>
__device__ __forceinline__ void sum_partials(const double* __restrict__ part, int count, double (&tot)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) tot[k] = 0;
  for (int b0 = threadIdx.x; b0 < count; b0 += U * NT) {
    double buf[U][K];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = b0 + u * NT;
#pragma unroll
      for (int k = 0; k < K; ++k) {  // a clamped index and a select: no branch around the load
        const double pv = part[(size_t)(b < count ? b : count - 1) * K + k];
        buf[u][k] = (b < count) ? pv : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int k = 0; k < K; ++k) tot[k] += buf[u][k];
  }
  block_sum<K, NT>(tot);
}

// sum_partials for count ≤ U·NT, as straight-line code: all U loads of a thread are issued before the first add
// (inside the general version's loop the scheduler may interleave them with waits).  Same order of additions.
template <int K, int NT, int U>
__device__ __forceinline__ void sum_partials_1pass(const double* __restrict__ part, int count, double (&tot)[K]) {
  double buf[U][K];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int b = (int)threadIdx.x + u * NT;
    const int bc = b < count ? b : count - 1;
#pragma unroll
    for (int k = 0; k < K; ++k) buf[u][k] = part[(size_t)bc * K + k];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) tot[k] = 0;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const bool in = (int)threadIdx.x + u * NT < count;
#pragma unroll
    for (int k = 0; k < K; ++k) tot[k] += in ? buf[u][k] : 0.0;
  }
  block_sum<K, NT>(tot);
}

// Item-major outputs with lane-contiguous stores.  Thread t of the workgroup holds the W values of item first + t
// (value k is get(k), k a compile-time index after unrolling); the `count` items that exist go to
// dst[(first + t)·W + k].  One lane per item would store with a stride of W doubles, each store instruction
// touching 64 cache lines; here the items pass through LDS one wave (64 items) at a time, at an odd stride so that
// the lane-strided LDS writes spread over the banks, and every thread then stores consecutive doubles of the
// wave's contiguous output span.  Every thread of the workgroup calls it (it synchronises the workgroup); lds
// holds 64·(W | 1) doubles.
template <int W, typename Get>
__device__ __forceinline__ void store_items(double* __restrict__ dst, long long first, int count, double* lds,
                                            Get&& get) {
  constexpr int S = W | 1;
  const int t = threadIdx.x;
#pragma unroll 1
  for (int g0 = 0; g0 < kBlock && g0 < count; g0 += 64) {  // uniform over the workgroup
    __syncthreads();  // the previous readers of lds are done
    if (t >= g0 && t < g0 + 64) {
#pragma unroll
      for (int k = 0; k < W; ++k) lds[(t - g0) * S + k] = get(k);
    }
    __syncthreads();
    const int n = (count - g0 < 64 ? count - g0 : 64) * W;
    double* __restrict__ out = dst + (first + g0) * W;
    for (int idx = t; idx < n; idx += kBlock) out[idx] = lds[(idx / W) * S + idx % W];
  }
}

// Host-side dispatch of the element kernels over (arithmetic type, material model): fn(ElemKind<T, Model>{}).
template <typename T, int Model>
struct ElemKind {
  using type = T;
  static constexpr int model = Model;
};

template <typename Fn>
void dispatch_elem(int model, bool fp32, Fn&& fn) {
  if (fp32) {
    if (model == kLinearElastic)
      fn(ElemKind<float, kLinearElastic>{});
    else
      fn(ElemKind<float, kStableNeoHookean>{});
  } else {
    if (model == kLinearElastic)
      fn(ElemKind<double, kLinearElastic>{});
    else
      fn(ElemKind<double, kStableNeoHookean>{});
  }
}

// ------------------------------------------------------------------------------------------------------
// element gathers

struct TetLocal {
  int id[4];
  double B[9];
  double V, lam, mu;
};

__device__ __forceinline__ TetLocal load_tet(const MeshView& m, int e) {
  TetLocal t;
  const int4 c = m.cells[e];
  t.id[0] = c.x;
  t.id[1] = c.y;
  t.id[2] = c.z;
  t.id[3] = c.w;
#pragma unroll
  for (int k = 0; k < 9; ++k) t.B[k] = m.B[(size_t)k * m.nt + e];
  t.V = m.V[e];
  t.lam = m.uniform ? m.lam0 : m.lam[e];  // uniform over the launch: no divergence
  t.mu = m.uniform ? m.mu0 : m.mu[e];
  return t;
}

// F = Ds·B with Ds from positions X + u (def_grad.hpp:270-288); differences formed in FP64
__device__ __forceinline__ void def_grad(const double x[4][3], const double B[9], double F[9]) {
  double Ds[9];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    Ds[a * 3 + 0] = x[1][a] - x[0][a];
    Ds[a * 3 + 1] = x[2][a] - x[0][a];
    Ds[a * 3 + 2] = x[3][a] - x[0][a];
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      F[a * 3 + c] = Ds[a * 3 + 0] * B[0 * 3 + c] + Ds[a * 3 + 1] * B[1 * 3 + c] + Ds[a * 3 + 2] * B[2 * 3 + c];
}

// local force f[v][b] = Σ_c P[b][c]·B[v−1][c], f[0] = −Σ f[1..3] (time_step.hpp:73-89)
template <typename T>
__device__ __forceinline__ void local_force(const T P[9], const T B[9], T f[4][3]) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    T tot = 0;
#pragma unroll
    for (int i = 1; i < 4; ++i) {
      const T val = P[j * 3 + 0] * B[(i - 1) * 3 + 0] + P[j * 3 + 1] * B[(i - 1) * 3 + 1] + P[j * 3 + 2] * B[(i - 1) * 3 + 2];
      f[i][j] = val;
      tot += val;
    }
    f[0][j] = -tot;
  }
}

// ------------------------------------------------------------------------------------------------------
// Hessian pullback

// Block (v, w), v ≤ w, of K = pfpxᵀ·H·pfpx: K_vw[b][d] = Σ_{c,e} D[v][c]·H[3b+c][3d+e]·D[w][e].
template <typename T, typename Emit>
__device__ __forceinline__ void pullback_blocks(const T H[45], const T D[4][3], Emit&& emit) {
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    T G[9][3];
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
      for (int dd = 0; dd < 3; ++dd)
        G[i][dd] = H[sym9(i, 3 * dd + 0)] * D[w][0] + H[sym9(i, 3 * dd + 1)] * D[w][1] + H[sym9(i, 3 * dd + 2)] * D[w][2];
#pragma unroll
    for (int v = 0; v <= w; ++v) {
      T K[9];
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int dd = 0; dd < 3; ++dd)
          K[b * 3 + dd] = D[v][0] * G[3 * b + 0][dd] + D[v][1] * G[3 * b + 1][dd] + D[v][2] * G[3 * b + 2][dd];
      emit(v, w, K);
    }
  }
}

template <typename T>
__device__ __forceinline__ void shape_grads(const T B[9], T D[4][3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    D[0][c] = -B[0 * 3 + c] - B[1 * 3 + c] - B[2 * 3 + c];
    D[1][c] = B[0 * 3 + c];
    D[2][c] = B[1 * 3 + c];
    D[3][c] = B[2 * 3 + c];
  }
}

// pair index of local blocks (v ≤ w): w(w+1)/2 + v
__host__ __device__ constexpr int pair_index(int v, int w) { return w * (w + 1) / 2 + v; }

// ------------------------------------------------------------------------------------------------------
// element bodies: energy + gradient (update_energy_and_gradients, time_step.hpp:577-624), element-wise:
//   e = ½·(ρV/20)·(Σ|Δ_i|² + |ΣΔ_i|²) + dt²·V·Ψ,   g_i = (ρV/20)·(Δ_i + ΣΔ_j) + dt²·V·f_i,   Δ = u − û;
// Hessian blocks (update_hessian, time_step.hpp:626-658; wrapped_pfpx :41-59).

// energy of tet e at deformation u (predictor uhat) and its 4 vertex gradient contributions c[i][a]
// With m.rest_x the rest shape (B = Dm⁻¹, V) is formed from the gathered rest positions by rest_shape, as the
// rest-shape kernel forms the stored B and V, instead of being loaded: the kernel then streams 16 B of cells (+ λ, μ
// when per tet) per tet instead of 96 B, for ~150 more FP64 operations per tet.
template <typename T, int Model>
__device__ __forceinline__ double elem_grad_contrib(const MeshView& m, int e, const double* __restrict__ u,
                                                    const double* __restrict__ uhat, double dt2, double c[4][3]) {
  TetLocal t;
  {
    const int4 ci = m.cells[e];
    t.id[0] = ci.x;
    t.id[1] = ci.y;
    t.id[2] = ci.z;
    t.id[3] = ci.w;
    t.lam = m.uniform ? m.lam0 : m.lam[e];  // uniform over the launch: no divergence
    t.mu = m.uniform ? m.mu0 : m.mu[e];
  }
  if (!m.rest_x) {  // uniform over the launch
#pragma unroll
    for (int k = 0; k < 9; ++k) t.B[k] = m.B[(size_t)k * m.nt + e];
    t.V = m.V[e];
  }
  double xr[4][3], x[4][3], d[4][3];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const size_t k = 3 * (size_t)t.id[i] + a;
      const double uu = u[k];
      xr[i][a] = m.X[k];
      x[i][a] = xr[i][a] + uu;
      d[i][a] = uu - uhat[k];
    }
  if (m.rest_x) rest_shape(xr[0], xr[1], xr[2], xr[3], t.B, &t.V);
  double Fd[9];
  def_grad(x, t.B, Fd);
  T F[9], Bt[9], P[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    F[k] = (T)Fd[k];
    Bt[k] = (T)t.B[k];
  }
  const T psi = material_stress<Model>(F, (T)t.lam, (T)t.mu, P);
  T f[4][3];
  local_force(P, Bt, f);
  const double mm = m.rho * t.V / 20.0;
  const double sc = dt2 * t.V;
  double S[3], dd = 0, ss = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) S[a] = d[0][a] + d[1][a] + d[2][a] + d[3][a];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      dd += d[i][a] * d[i][a];
      c[i][a] = mm * (d[i][a] + S[a]) + sc * (double)f[i][a];
    }
#pragma unroll
  for (int a = 0; a < 3; ++a) ss += S[a] * S[a];
  return 0.5 * mm * (dd + ss) + sc * (double)psi;
}

// Tile phase of the tile-reduced gradient (every thread of the TILE-thread workgroup calls it; it synchronises):
// thread t evaluates tet e = tile·TILE + t when `active`, puts its 4 vertex contributions into LDS at the tile-local slots
// tslot[e] (the tile's corners sorted by (vertex, tet, corner)), and the workgroup then sums each tile vertex's
// contiguous run in slot order into gpart[gt_off[tile] + j].  Returns the thread's element energy.
template <typename T, int Model, int TILE = kBlock>
__device__ __forceinline__ double elem_grad_tile(const MeshView& m, int tile, bool active, const double* __restrict__ u,
                                                 const double* __restrict__ uhat, double dt2,
                                                 const ushort4* __restrict__ tslot, const int* __restrict__ gt_off,
                                                 const int* __restrict__ gt_run, double* __restrict__ gpart) {
  __shared__ double rec[4 * TILE * 3];
  const int e = tile * TILE + (int)threadIdx.x;
  double energy = 0.0;
  if (active) {
    double c[4][3];
    energy = elem_grad_contrib<T, Model>(m, e, u, uhat, dt2, c);
    const ushort4 sl = tslot[e];
    const int slot[4] = {sl.x, sl.y, sl.z, sl.w};
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int a = 0; a < 3; ++a) rec[3 * slot[i] + a] = c[i][a];
  }
  __syncthreads();
  const int t0 = gt_off[tile], ntv = gt_off[tile + 1] - t0;
  const int* __restrict__ run = gt_run + t0 + tile;  // ntv + 1 run boundaries of this tile
  for (int j = threadIdx.x; j < ntv; j += TILE) {
    double acc[3] = {0.0, 0.0, 0.0};
    const int q1 = run[j + 1];
    for (int q = run[j]; q < q1; ++q)
#pragma unroll
      for (int a = 0; a < 3; ++a) acc[a] += rec[3 * q + a];
#pragma unroll
    for (int a = 0; a < 3; ++a) gpart[3 * (size_t)(t0 + j) + a] = acc[a];
  }
  return energy;
}

// g[v] = Σ of v's tile partials in tile order, zeroed on Dirichlet dofs (boundary.hpp:45-51); returns v's share
// of ‖g‖²
__device__ __forceinline__ double grad_gather_vertex(int v, const int* __restrict__ gv_seg,
                                                     const int* __restrict__ gv_pos, const double* __restrict__ gpart,
                                                     const unsigned char* __restrict__ dof, double* __restrict__ g) {
  double acc[3] = {0.0, 0.0, 0.0};
  const int k1 = gv_seg[v + 1];
  for (int k = gv_seg[v]; k < k1; ++k) {
    const size_t p = 3 * (size_t)gv_pos[k];
#pragma unroll
    for (int a = 0; a < 3; ++a) acc[a] += gpart[p + a];
  }
  double gg = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double val = dof[3 * (size_t)v + a] ? 0.0 : acc[a];
    g[3 * (size_t)v + a] = val;
    gg += val * val;
  }
  return gg;
}

// the 10 oriented upper blocks of tet e (dt²V-scaled stiffness plus consistent mass) to the slots hperm[10e+k]
template <typename T, int Model>
__device__ __forceinline__ void elem_hess_tet(const MeshView& m, int e, const double* __restrict__ u, double dt2,
                                              int make_spsd, const int* __restrict__ hperm,
                                              double* __restrict__ hcontrib) {
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
  T F[9], Bt[9], H[45], D[4][3];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    F[k] = (T)Fd[k];
    Bt[k] = (T)t.B[k];
  }
  material_hessian<Model>(F, (T)t.lam, (T)t.mu, make_spsd, (T)kSpdFloor, H);
  shape_grads(Bt, D);
  const double sc = dt2 * t.V;
  const double mm = m.rho * t.V / 20.0;
  const int* pp = hperm + 10 * (size_t)e;
  pullback_blocks(H, D, [&](int v, int w, const T K[9]) {
    const size_t pos = 9 * (size_t)pp[pair_index(v, w)];
    const double mvw = (v == w) ? 2.0 * mm : mm;
    const bool tr = t.id[v] > t.id[w];
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int dd = 0; dd < 3; ++dd) {
        const double kv = sc * (double)(tr ? K[dd * 3 + b] : K[b * 3 + dd]) + (b == dd ? mvw : 0.0);
        hcontrib[pos + b * 3 + dd] = kv;
      }
  });
}

// Compare-exchange of two corners by vertex id (with their shape-gradient rows); compile-time indices only.
template <typename T, int A, int B>
__device__ __forceinline__ void corner_cx(int (&id)[4], T (&D)[4][3]) {
  const bool sw = id[A] > id[B];
  const int ia = id[A], ib = id[B];
  id[A] = sw ? ib : ia;
  id[B] = sw ? ia : ib;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T da = D[A][c], db = D[B][c];
    D[A][c] = sw ? db : da;
    D[B][c] = sw ? da : db;
  }
}

// Tile phase of the tile-reduced Hessian assembly (every thread of the TILE-thread workgroup calls it; it
// synchronises).  Thread t takes tet e = tile·TILE + t when `active`: its 9x9 material Hessian H (SPD-projected when
// make_spsd) and shape gradients D, with the four corners put in ascending vertex order, so that local block (v, w),
// v ≤ w in that order, is the upper block (min id, max id) of the global matrix (K_wv = K_vwᵀ since H is
// symmetric).  For each block row i it forms the components (i, 0..2) of its 10 blocks
// K_vw = dt²V·D_vᵀ·H·D_w plus the consistent mass ρV(1+δ_vw)/20 on the diagonal components, puts them into LDS at
// its tile-local slots (hslot: the tile's contributions sorted by (upper block, tet, pair), two 16-bit slots per
// word, word j of tet e at hslot[j·nt + e]), and the workgroup sums every tile block's contiguous run in slot order
// into one partial per (tile, block) and component: hpart[comp·P + ht_off[tile] + j].  Only those partials reach
// HBM (about 2.8 blocks per tet on a structured bar with 256-tet tiles instead of 10).
template <typename T, int Model, int TILE>
__device__ __forceinline__ void elem_hess_tile(const MeshView& m, int tile, bool active, const double* __restrict__ u,
                                               double dt2, int make_spsd, const unsigned* __restrict__ hslot,
                                               const int* __restrict__ ht_off, const int* __restrict__ ht_run,
                                               double* __restrict__ hpart, long long P) {
  constexpr int S = 10 * TILE;
  __shared__ double rec[3 * S];
  const int e = tile * TILE + (int)threadIdx.x;
  T H[45], D[4][3];
  double sc = 0.0, mm = 0.0;
  unsigned sw[5] = {0u, 0u, 0u, 0u, 0u};
  if (active) {
    TetLocal t;
    {
      const int4 ci = m.cells[e];
      t.id[0] = ci.x;
      t.id[1] = ci.y;
      t.id[2] = ci.z;
      t.id[3] = ci.w;
      t.lam = m.uniform ? m.lam0 : m.lam[e];  // uniform over the launch: no divergence
      t.mu = m.uniform ? m.mu0 : m.mu[e];
    }
    if (!m.rest_x) {  // uniform over the launch
#pragma unroll
      for (int k = 0; k < 9; ++k) t.B[k] = m.B[(size_t)k * m.nt + e];
      t.V = m.V[e];
    }
#pragma unroll
    for (int j = 0; j < 5; ++j) sw[j] = hslot[(size_t)j * m.nt + e];
    double xr[4][3], x[4][3];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const size_t k = 3 * (size_t)t.id[i] + a;
        xr[i][a] = m.X[k];
        x[i][a] = xr[i][a] + u[k];
      }
    if (m.rest_x) rest_shape(xr[0], xr[1], xr[2], xr[3], t.B, &t.V);
    double Fd[9];
    def_grad(x, t.B, Fd);
    T F[9], Bt[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      F[k] = (T)Fd[k];
      Bt[k] = (T)t.B[k];
    }
    material_hessian<Model>(F, (T)t.lam, (T)t.mu, make_spsd, (T)kSpdFloor, H);
    shape_grads(Bt, D);
    int id[4] = {t.id[0], t.id[1], t.id[2], t.id[3]};
    corner_cx<T, 0, 1>(id, D);
    corner_cx<T, 2, 3>(id, D);
    corner_cx<T, 0, 2>(id, D);
    corner_cx<T, 1, 3>(id, D);
    corner_cx<T, 1, 2>(id, D);
    sc = dt2 * t.V;
    mm = m.rho * t.V / 20.0;
  }
  const int t0 = ht_off[tile], nb = ht_off[tile + 1] - t0;
  const int* __restrict__ run = ht_run + t0 + tile;  // nb + 1 run boundaries of this tile
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (active) {
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) {
        T G[4][3];
#pragma unroll
        for (int w = 0; w < 4; ++w)
#pragma unroll
          for (int c = 0; c < 3; ++c)
            G[w][c] = H[sym9(3 * i + c, 3 * jj + 0)] * D[w][0] + H[sym9(3 * i + c, 3 * jj + 1)] * D[w][1] +
                      H[sym9(3 * i + c, 3 * jj + 2)] * D[w][2];
#pragma unroll
        for (int w = 0; w < 4; ++w)
#pragma unroll
          for (int v = 0; v <= w; ++v) {
            const int k = pair_index(v, w);
            double val = sc * (double)(D[v][0] * G[w][0] + D[v][1] * G[w][1] + D[v][2] * G[w][2]);
            if (i == jj) val += (v == w) ? 2.0 * mm : mm;
            rec[jj * S + (int)((sw[k >> 1] >> (16 * (k & 1))) & 0xffffu)] = val;
          }
      }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < nb; j += TILE) {
      double a0 = 0.0, a1 = 0.0, a2 = 0.0;
      const int q1 = run[j + 1];
      for (int q = run[j]; q < q1; ++q) {
        a0 += rec[q];
        a1 += rec[S + q];
        a2 += rec[2 * S + q];
      }
      double* __restrict__ dst = hpart + (size_t)(3 * i) * P + t0 + j;
      dst[0] = a0;
      dst[P] = a1;
      dst[2 * P] = a2;
    }
    __syncthreads();
  }
}

// upper block ub: the sum of its tile partials in tile order (hb_pos[hb_seg[ub] .. hb_seg[ub+1])).  The first four
// partials' loads are issued together (clamped positions, no branch; a block of a structured mesh has at most four),
// then added in order; longer runs continue one partial at a time.
__device__ __forceinline__ void hess_gather_block(int ub, const int* __restrict__ hb_seg,
                                                  const int* __restrict__ hb_pos, const double* __restrict__ hpart,
                                                  long long P, double (&S)[9]) {
  constexpr int U = 4;
  const int k0 = hb_seg[ub], k1 = hb_seg[ub + 1];
#pragma unroll
  for (int c = 0; c < 9; ++c) S[c] = 0.0;
  if (k1 <= k0) return;
  size_t p[U];
#pragma unroll
  for (int i = 0; i < U; ++i) p[i] = (size_t)hb_pos[min(k0 + i, k1 - 1)];
  double v[U][9];
#pragma unroll
  for (int i = 0; i < U; ++i)
#pragma unroll
    for (int c = 0; c < 9; ++c) v[i][c] = hpart[(size_t)c * P + p[i]];
#pragma unroll
  for (int i = 0; i < U; ++i)
#pragma unroll
    for (int c = 0; c < 9; ++c) S[c] += (k0 + i < k1) ? v[i][c] : 0.0;
  for (int k = k0 + U; k < k1; ++k) {
    const size_t q = (size_t)hb_pos[k];
#pragma unroll
    for (int c = 0; c < 9; ++c) S[c] += hpart[(size_t)c * P + q];
  }
}

// Element part of the adjoint of one implicit step, for the adjoint vector l (solution of H l = ∂L/∂u at the
// converged u):
//   mass contributions (ρV/20)(l_i + Σ_j l_j) to the gradient slots gperm[4e+i] (reduced per vertex: M l),
//   ∂L/∂λ_e = −dt²·V·(∂P/∂λ : G),  ∂L/∂μ_e = −dt²·V·(∂P/∂μ : G),  G = Ds(l)·B = Σ_i l_i ⊗ ∇N_i,
// because the step's residual depends on (λ, μ) only through the elastic forces V·P·∇N_i.  With
// P = λ(J − α)·cof F + μ(1 − 1/(Ic+δ))·F and α = 1 + (1 − 1/(3+δ))·μ/λ (elast/stable_neohookean.hpp:23-62):
//   ∂P/∂λ = (J − 1)·cof F,   ∂P/∂μ = −(1 − 1/(3+δ))·cof F + (1 − 1/(Ic+δ))·F.
// With the linear model's P = 2μ·ε + λ·tr(ε)·I (elast/linear.hpp:37-50): ∂P/∂λ = tr(ε)·I, ∂P/∂μ = 2ε.
template <int Model>
__device__ __forceinline__ void elem_adjoint_tet(const MeshView& m, int e, const double* __restrict__ u,
                                                 const double* __restrict__ lv, double dt2,
                                                 const int* __restrict__ gperm, double* __restrict__ gcontrib,
                                                 double* __restrict__ g_lam, double* __restrict__ g_mu) {
  const TetLocal t = load_tet(m, e);
  double x[4][3], l[4][3];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const size_t k = 3 * (size_t)t.id[i] + a;
      x[i][a] = m.X[k] + u[k];
      l[i][a] = lv[k];
    }
  double F[9], G[9];
  def_grad(x, t.B, F);
  def_grad(l, t.B, G);
  const double sc = -dt2 * t.V;
  if constexpr (Model == kLinearElastic) {
    const double tr = F[0] + F[4] + F[8] - 3.0;
    double eG = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        eG += (0.5 * (F[3 * a + c] + F[3 * c + a]) - (a == c ? 1.0 : 0.0)) * G[3 * a + c];
    g_lam[e] = sc * tr * (G[0] + G[4] + G[8]);
    g_mu[e] = sc * 2.0 * eG;
  } else {
    double C[9];
    cofactor(F, C);
    double Ic = 0, cG = 0, fG = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      Ic += F[k] * F[k];
      cG += C[k] * G[k];
      fG += F[k] * G[k];
    }
    const double J = det3(F);
    constexpr double del = 1.0;
    g_lam[e] = sc * (J - 1.0) * cG;
    g_mu[e] = sc * (-(1.0 - 1.0 / (3.0 + del)) * cG + (1.0 - 1.0 / (Ic + del)) * fG);
  }
  const double mm = m.rho * t.V / 20.0;
  double S[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) S[a] = l[0][a] + l[1][a] + l[2][a] + l[3][a];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const size_t pos = 3 * (size_t)gperm[4 * (size_t)e + i];
#pragma unroll
    for (int a = 0; a < 3; ++a) gcontrib[pos + a] = mm * (l[i][a] + S[a]);
  }
}

// LDS-staged segmented sums.  Thread t of the workgroup owns segment j = j0 + t (j < j1), the record run
// [seg[j], seg[j+1]) of W doubles per record.  The runs of consecutive segments are contiguous, so the workgroup
// copies its whole span [seg[j0], seg[j1]) through LDS with coalesced loads, kStageDoubles at a time (all of a
// stage's loads in flight together), and each thread adds its own records in index order: the sums are bitwise
// those of a sequential walk over the run, but the loads are no longer one scattered walk per lane.  Every thread
// of the workgroup must call it (it synchronises the workgroup); threads with j ≥ j1 get zeros.
constexpr int kStageDoubles = 4096;  // 32 KB of LDS per workgroup

template <int W>
__device__ __forceinline__ void staged_run_sum(const int* __restrict__ seg, int j0, int j1,
                                               const double* __restrict__ rec, double (&acc)[W]) {
  constexpr int kStage = kStageDoubles / W;  // records per stage
  constexpr int U = kStageDoubles / kBlock;  // loads in flight per thread: a whole stage at once
  __shared__ double lds[kStage * W];
  const int j = j0 + (int)threadIdx.x;
  const int span_lo = seg[j0], span_hi = seg[j1];
  int lo = 0, hi = 0;
  if (j < j1) {
    lo = seg[j];
    hi = seg[j + 1];
  }
#pragma unroll
  for (int k = 0; k < W; ++k) acc[k] = 0.0;
  for (int s0 = span_lo; s0 < span_hi; s0 += kStage) {
    const int cnt = min(kStage, span_hi - s0) * W;
    const double* __restrict__ src = rec + (size_t)s0 * W;
    for (int t0 = 0; t0 < cnt; t0 += U * kBlock) {
      double buf[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int t = t0 + u * kBlock + (int)threadIdx.x;
        buf[u] = t < cnt ? src[t] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int t = t0 + u * kBlock + (int)threadIdx.x;
        if (t < cnt) lds[t] = buf[u];
      }
    }
    __syncthreads();
    const int a = max(lo, s0) - s0, b = min(hi, s0 + kStage) - s0;
    for (int q = a; q < b; ++q)
#pragma unroll
      for (int k = 0; k < W; ++k) acc[k] += lds[q * W + k];
    __syncthreads();
  }
}

// segmented reduce of the gradient contributions of vertices [v0, v1), one per thread, zeroed on Dirichlet dofs
// (boundary.hpp:45-51); returns the thread's vertex's share of ‖g‖²
__device__ __forceinline__ double grad_reduce_vertices(int v0, int v1, const int* __restrict__ gseg,
                                                       const double* __restrict__ gcontrib,
                                                       const unsigned char* __restrict__ dof,
                                                       double* __restrict__ g) {
  double acc[3];
  staged_run_sum<3>(gseg, v0, v1, gcontrib, acc);
  const int v = v0 + (int)threadIdx.x;
  double gg = 0.0;
  if (v < v1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double val = dof[3 * (size_t)v + a] ? 0.0 : acc[a];
      g[3 * (size_t)v + a] = val;
      gg += val * val;
    }
  }
  return gg;
}

// upper block ub = (rc.x, rc.y) with its summed contributions S: Dirichlet identity filter, store into the
// blocked-ELL values; Jacobi inverse diagonal (and optionally the diagonal) for a diagonal block.
// ub_pos[ub] ≥ 0 is the value base of a 9-value block; a negative entry −(base + 1) marks a diagonal block of
// the fixed layout, which keeps the upper triangle (00, 01, 02, 11, 12, 22) only.
__device__ __forceinline__ void hess_store_block(int ub, int2 rc, double (&S)[9], const int* __restrict__ ub_pos,
                                                 const unsigned char* __restrict__ dof, int filter_dbc,
                                                 double* __restrict__ val, double* __restrict__ dinv,
                                                 double* __restrict__ diag) {
  if (filter_dbc) {
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int dd = 0; dd < 3; ++dd)
        if (dof[3 * (size_t)rc.x + b] || dof[3 * (size_t)rc.y + dd]) S[b * 3 + dd] = (rc.x == rc.y && b == dd) ? 1.0 : 0.0;
  }
  const int praw = ub_pos[ub];
  if (praw >= 0) {
    const size_t pos = (size_t)praw;
#pragma unroll
    for (int k = 0; k < 9; ++k) val[pos + (size_t)k * kSlice] = S[k];
  } else {
    const size_t pos = (size_t)(-(praw + 1));
    constexpr int kUpper[6] = {0, 1, 2, 4, 5, 8};
#pragma unroll
    for (int m = 0; m < 6; ++m) val[pos + (size_t)m * kSlice] = S[kUpper[m]];
  }
  if (rc.x == rc.y) {
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      dinv[3 * (size_t)rc.x + b] = 1.0 / S[b * 4];
      if (diag) diag[3 * (size_t)rc.x + b] = S[b * 4];
    }
  }
}

// ------------------------------------------------------------------------------------------------------
// blocked-ELL row product

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X.  With remapping, physical workgroup i
// handles logical block (i mod 8)·per + i/8, so each XCD (and its L2) sees a contiguous range of rows and the
// transposed reads of a row's lower blocks mostly hit lines its own XCD has just read.  Returns −1 for the idle
// tail of the padded grid.
__device__ __forceinline__ int logical_block(int nb, int remap) {
  if (!remap) return blockIdx.x;
  const int per = (nb + 7) >> 3;
  const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  return lb < nb ? lb : -1;
}

inline int grid_for(int nb, int remap) { return remap ? ((nb + 7) / 8) * 8 : nb; }

// SpMV workgroups are single waves: one 64-row slice per workgroup (LPR ≤ 1) spreads the ~2,200 waves of the
// 786k-tet bar evenly over the 1,024 SIMDs in one dispatch round.
constexpr int kSpmvBlock = 64;

template <int LPR>
struct RowMap {
  static constexpr int rows_per_block = kSpmvBlock / (LPR > 0 ? LPR : 1);
  // row and slot phase of this thread within logical block lb
  __device__ static __forceinline__ void map(int lb, int& v, int& j) {
    if (LPR <= 1) {
      v = lb * kSpmvBlock + threadIdx.x;
      j = 0;
    } else {
      const int lane = threadIdx.x & 63;
      v = lb * rows_per_block + (threadIdx.x >> 6) * 16 + (lane & 15);
      j = lane >> 4;
    }
  }
  __host__ __device__ static int blocks(int nrows) { return (nrows + rows_per_block - 1) / rows_per_block; }
};

template <int LPR>
__device__ __forceinline__ void bell_row_part(const BellView& A, int v, int j, const double* __restrict__ x,
                                              double y[3]) {
  const int s = v >> 6, lane = v & 63;
  double y0 = 0, y1 = 0, y2 = 0;
  {
    const int W = A.wu[s];
    const int off = A.offu[s];
    const int* col = A.ucol + off + lane;
    const double* val = A.val + 9 * (size_t)off + lane;
#pragma unroll 4
    for (int k = j; k < W; k += LPR) {
      const size_t w3 = 3 * (size_t)col[k * kSlice];
      const double x0 = x[w3], x1 = x[w3 + 1], x2 = x[w3 + 2];
      const double* vk = val + (size_t)k * 9 * kSlice;
      y0 += vk[0 * kSlice] * x0 + vk[1 * kSlice] * x1 + vk[2 * kSlice] * x2;
      y1 += vk[3 * kSlice] * x0 + vk[4 * kSlice] * x1 + vk[5 * kSlice] * x2;
      y2 += vk[6 * kSlice] * x0 + vk[7 * kSlice] * x1 + vk[8 * kSlice] * x2;
    }
  }
  {
    const int W = A.wl[s];
    const int2* ref = A.lref + A.offl[s] + lane;
#pragma unroll 4
    for (int k = j; k < W; k += LPR) {
      const int2 rw = ref[k * kSlice];
      const size_t w3 = 3 * (size_t)rw.x;
      const double x0 = x[w3], x1 = x[w3 + 1], x2 = x[w3 + 2];
      const double* b = A.val + (size_t)rw.y;
      y0 += b[0 * kSlice] * x0 + b[3 * kSlice] * x1 + b[6 * kSlice] * x2;
      y1 += b[1 * kSlice] * x0 + b[4 * kSlice] * x1 + b[7 * kSlice] * x2;
      y2 += b[2 * kSlice] * x0 + b[5 * kSlice] * x1 + b[8 * kSlice] * x2;
    }
  }
  y[0] = y0;
  y[1] = y1;
  y[2] = y2;
}

// Row product over the fixed layout (every row padded to 7 upper and 7 lower off-diagonal slots plus the 6-value
// diagonal, 16-bit indices): straight-line code, the row's two 16-byte index words loaded first.  The slots are
// aligned by direction where the mesh allows (lower slot 7 − k refers to blocks stored in upper slot k of the
// neighbour rows), and each pair (lower 7 − k, upper k) is issued together, the far directions first: the lower
// reads of a slice then reach L2 at about the time the slice that stores those blocks reads them itself.
struct FixedRowIndex {
  uint4 i0, i1;
};

__device__ __forceinline__ FixedRowIndex bell_fixed_index(const BellView& A, int v) {
  static_assert(kFixedWU == 8 && kFixedWL == 7, "two 16-byte index words per row");
  const int s = v >> 6, lane = v & 63;
  return {A.idx[(2 * s) * kSlice + lane], A.idx[(2 * s + 1) * kSlice + lane]};
}

// the row product given the row's index words (loaded earlier by the caller, e.g. before a wait on scalars)
__device__ __forceinline__ void bell_row_fixed(const BellView& A, int v, const FixedRowIndex& ix,
                                               const double* __restrict__ x, double y[3]) {
  const int s = v >> 6, lane = v & 63;
  const uint4 i0 = ix.i0, i1 = ix.i1;
  const unsigned up[7] = {i0.x & 0xffffu, i0.x >> 16, i0.y & 0xffffu, i0.y >> 16,
                          i0.z & 0xffffu, i0.z >> 16, i0.w & 0xffffu};
  const unsigned lo[7] = {i0.w >> 16, i1.x & 0xffffu, i1.x >> 16, i1.y & 0xffffu,
                          i1.y >> 16,  i1.z & 0xffffu, i1.z >> 16};
  const unsigned slots = i1.w;
  const double* val = A.val + kBellZeroLead + (size_t)s * kBellSliceVals + lane;
  double y0 = 0, y1 = 0, y2 = 0;
  auto upper = [&](int k) {  // k = 1..7
    const size_t w3 = 3 * (size_t)(v + (int)up[k - 1]);
    const double x0 = x[w3], x1 = x[w3 + 1], x2 = x[w3 + 2];
    const double* vk = val + (size_t)(6 + 9 * (k - 1)) * kSlice;
    y0 += vk[0 * kSlice] * x0 + vk[1 * kSlice] * x1 + vk[2 * kSlice] * x2;
    y1 += vk[3 * kSlice] * x0 + vk[4 * kSlice] * x1 + vk[5 * kSlice] * x2;
    y2 += vk[6 * kSlice] * x0 + vk[7 * kSlice] * x1 + vk[8 * kSlice] * x2;
  };
  auto lower = [&](int j) {  // j = 0..6
    const int w = v - (int)lo[j];
    const int kw = (int)((slots >> (3 * j)) & 7u);
    const size_t w3 = 3 * (size_t)w;
    const double x0 = x[w3], x1 = x[w3 + 1], x2 = x[w3 + 2];
    // value offsets fit 32 bits (checked when the layout is built); padding (kw = 0) reads the leading zero
    // block.  Selected with a mask: a conditional expression here becomes a branch per block.
    const unsigned real = (unsigned)kBellZeroLead + (unsigned)(w >> 6) * (unsigned)kBellSliceVals +
                          (unsigned)(6 + 9 * (kw - 1)) * kSlice + (unsigned)(w & 63);
    const unsigned m = 0u - (unsigned)(kw != 0);
    const double* b = A.val + ((real & m) | ((unsigned)lane & ~m));
    y0 += b[0 * kSlice] * x0 + b[3 * kSlice] * x1 + b[6 * kSlice] * x2;
    y1 += b[1 * kSlice] * x0 + b[4 * kSlice] * x1 + b[7 * kSlice] * x2;
    y2 += b[2 * kSlice] * x0 + b[5 * kSlice] * x1 + b[8 * kSlice] * x2;
  };
#pragma unroll
  for (int k = kFixedWU - 1; k >= 1; --k) {
    lower(kFixedWU - 1 - k);
    upper(k);
  }
  {
    const size_t v3 = 3 * (size_t)v;
    const double x0 = x[v3], x1 = x[v3 + 1], x2 = x[v3 + 2];
    const double d00 = val[0 * kSlice], d01 = val[1 * kSlice], d02 = val[2 * kSlice];
    const double d11 = val[3 * kSlice], d12 = val[4 * kSlice], d22 = val[5 * kSlice];
    y0 += d00 * x0 + d01 * x1 + d02 * x2;
    y1 += d01 * x0 + d11 * x1 + d12 * x2;
    y2 += d02 * x0 + d12 * x1 + d22 * x2;
  }
  y[0] = y0;
  y[1] = y1;
  y[2] = y2;
}

// LPR = 0 selects the fixed-layout single-lane row product
template <int LPR>
__device__ __forceinline__ void row_part(const BellView& A, int v, int j, const double* __restrict__ x, double y[3]) {
  if (LPR == 0)
    bell_row_fixed(A, v, bell_fixed_index(A, v), x, y);
  else
    bell_row_part<LPR>(A, v, j, x, y);
}

// sum over the LPR lanes of a row (lanes l, l+16, l+32, l+48 for LPR = 4); every lane gets the total
template <int LPR>
__device__ __forceinline__ void row_combine(double y[3]) {
  if (LPR <= 1) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    y[a] += __shfl_xor(y[a], 16, 64);
    y[a] += __shfl_xor(y[a], 32, 64);
  }
}

}  // namespace ssim_hip
