// batch_kernels.hip -- kernels of batched stepping: many meshes of uneven size advanced by one set of launches.
//
// The batch is one concatenated mesh whose per-mesh vertex and tet ranges are padded to whole workgroups, so
// a workgroup reads its mesh from a small table and every per-workgroup partial sum belongs to one mesh.  The
// element bodies, the segmented reduces and the blocked-ELL row product are the single-mesh ones; what changes
// is that scalars (energies, dot products, PCG ρ, α, β, convergence) are per mesh, and a workgroup whose mesh
// is masked out or has converged returns at entry.
#include <hip/hip_runtime.h>

#include "device_common.h"

namespace ssim_hip {

namespace {

// dofs of vertex workgroup w handled by thread t: 3·kBatchRows·w + t + kBatchRows·j, j = 0..2 (coalesced)
constexpr int kBVec = 3;

__device__ __forceinline__ size_t bdof(int j) {
  return (size_t)blockIdx.x * (kBVec * kBatchRows) + (size_t)j * kBatchRows + threadIdx.x;
}

// the same for an explicit vertex workgroup w (compacted grids)
__device__ __forceinline__ size_t bdof_w(int w, int j) {
  return (size_t)w * (kBVec * kBatchRows) + (size_t)j * kBatchRows + threadIdx.x;
}

// The totals of two width-2 partial arrays over the same `count` entries, with the same additions as two
// sum_partials<2, NT, 1> calls (each lane adds its entries in index order, then the workgroup sum), but with a
// trip's loads from both arrays in flight together: one memory round instead of two.
template <int NT>
__device__ __forceinline__ void sum_partials_pair(const double* __restrict__ a, const double* __restrict__ b,
                                                  int count, double (&ta)[2], double (&tb)[2]) {
  double t[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < count; i += NT) {
    const double a0 = a[2 * (size_t)i], a1 = a[2 * (size_t)i + 1];
    const double b0 = b[2 * (size_t)i], b1 = b[2 * (size_t)i + 1];
    t[0] += a0;
    t[1] += a1;
    t[2] += b0;
    t[3] += b1;
  }
  block_sum<4, NT>(t);
  ta[0] = t[0];
  ta[1] = t[1];
  tb[0] = t[2];
  tb[1] = t[3];
}

// The total of `na` width-1 partials `a` and of the first nb ≤ na width-2 partials `b`, with the same additions as
// sum_partials<1, kBlock, 1>(a, na) and sum_partials<2, kBlock, 1>(b, nb), in one loop and one workgroup sum.
__device__ __forceinline__ void sum_partials_1_2(const double* __restrict__ a, int na, const double* __restrict__ b,
                                                 int nb, double (&ta)[1], double (&tb)[2]) {
  double t[3] = {0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < na; i += kBlock) {
    const double av = a[i];
    const bool in = i < nb;
    const size_t j = 2 * (size_t)(in ? i : 0);
    const double b0 = b[j], b1 = b[j + 1];
    t[0] += av;
    if (in) {
      t[1] += b0;
      t[2] += b1;
    }
  }
  block_sum<3, kBlock>(t);
  ta[0] = t[0];
  tb[0] = t[1];
  tb[1] = t[2];
}

// ------------------------------------------------------------------------------------------------------
// energy + gradient

__global__ __launch_bounds__(kBlock) void k_bgrad_reduce(BatchView bt, const int* __restrict__ gseg,
                                                         const double* __restrict__ gcontrib,
                                                         const unsigned char* __restrict__ dof, double* __restrict__ g,
                                                         double* __restrict__ part_g) {
  if (!bt.mask[bt.vwg_mesh[blockIdx.x]]) return;  // uniform over the workgroup
  const int v0 = blockIdx.x * kBlock;
  double gg[1] = {grad_reduce_vertices(v0, v0 + kBlock, gseg, gcontrib, dof, g)};
  store_partials<1>(gg, part_g);
}

// the tile-reduced evaluation (single-mesh k_elem_grad_tile): tet workgroup b is tile b of the concatenated mesh,
// whose tets and vertices all belong to one mesh
template <typename T, int Model>
__global__ __launch_bounds__(kBlock) void k_belem_grad_tile(MeshView m, BatchView bt, const double* __restrict__ u,
                                                            const double* __restrict__ uhat, double dt2,
                                                            const ushort4* __restrict__ tslot,
                                                            const int* __restrict__ gt_off,
                                                            const int* __restrict__ gt_run, double* __restrict__ gpart,
                                                            double* __restrict__ part_e) {
  const int mesh = bt.twg_mesh[blockIdx.x];
  if (!bt.mask[mesh]) return;  // uniform over the workgroup
  const int e = blockIdx.x * kBlock + threadIdx.x;
  double energy[1] = {elem_grad_tile<T, Model>(m, blockIdx.x, e < bt.tend[mesh], u, uhat, dt2, tslot, gt_off, gt_run,
                                               gpart)};
  store_partials<1>(energy, part_e);
}

__global__ __launch_bounds__(kBlock) void k_bgrad_gather(BatchView bt, const int* __restrict__ gv_seg,
                                                         const int* __restrict__ gv_pos,
                                                         const double* __restrict__ gpart,
                                                         const unsigned char* __restrict__ dof, double* __restrict__ g,
                                                         double* __restrict__ part_g) {
  if (!bt.mask[bt.vwg_mesh[blockIdx.x]]) return;  // uniform over the workgroup
  double gg[1] = {grad_gather_vertex(blockIdx.x * kBlock + threadIdx.x, gv_seg, gv_pos, gpart, dof, g)};
  store_partials<1>(gg, part_g);
}

// one workgroup per mesh: energy over the mesh's tet workgroups, ‖g‖² over its vertex workgroups
__global__ __launch_bounds__(kBlock) void k_bfinalize_eval(BatchView bt, const double* __restrict__ part_e,
                                                           const double* __restrict__ part_g, EvalScalars* out) {
  const int mesh = blockIdx.x;
  if (!bt.mask[mesh]) return;
  double te[1], tg[1] = {0.0};
  const int t0 = bt.twg_off[mesh], v0 = bt.vwg_off[mesh];
  sum_partials<1>(part_e + t0, bt.twg_off[mesh + 1] - t0, te);
  if (part_g) sum_partials<1>(part_g + v0, bt.vwg_off[mesh + 1] - v0, tg);
  if (threadIdx.x == 0) {
    out[mesh].energy = te[0];
    if (part_g) out[mesh].gg = tg[0];
  }
}

__global__ __launch_bounds__(kBlock) void k_bdot_part(BatchView bt, const double* __restrict__ a,
                                                      const double* __restrict__ b, double* __restrict__ part) {
  if (!bt.mask[bt.vwg_mesh[blockIdx.x]]) return;
  double v[1] = {0.0};
#pragma unroll
  for (int j = 0; j < kBVec; ++j) {
    const size_t i = bdof(j);
    v[0] += a[i] * b[i];
  }
  store_partials<1>(v, part);
}

__global__ __launch_bounds__(kBlock) void k_bfinalize_dot(BatchView bt, const double* __restrict__ part,
                                                          EvalScalars* out) {
  const int mesh = blockIdx.x;
  if (!bt.mask[mesh]) return;
  double t[1];
  const int v0 = bt.vwg_off[mesh];
  sum_partials<1>(part + v0, bt.vwg_off[mesh + 1] - v0, t);
  if (threadIdx.x == 0) out[mesh].gd = t[0];
}

// ------------------------------------------------------------------------------------------------------
// Hessian

template <typename T, int Model>
__global__ __launch_bounds__(kBlock) void k_belem_hess(MeshView m, BatchView bt, const double* __restrict__ u,
                                                       double dt2, int make_spsd, const int* __restrict__ hperm,
                                                       double* __restrict__ hcontrib) {
  const int mesh = bt.twg_mesh[blockIdx.x];
  if (!bt.mask[mesh]) return;
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= bt.tend[mesh]) return;
  elem_hess_tet<T, Model>(m, e, u, dt2, make_spsd, hperm, hcontrib);
}

// tile-reduced assembly (single-mesh k_elem_hess_tile): tet workgroup b is tile b of the concatenated mesh, whose
// per-mesh tet ranges are padded to whole tiles, so a tile belongs to one mesh and a masked-out mesh skips it whole
template <typename T, int Model>
__global__ __launch_bounds__(kHessTile) void k_belem_hess_tile(MeshView m, BatchView bt, const double* __restrict__ u,
                                                               double dt2, int make_spsd,
                                                               const unsigned* __restrict__ hslot,
                                                               const int* __restrict__ ht_off,
                                                               const int* __restrict__ ht_run,
                                                               double* __restrict__ hpart, long long P) {
  const int mesh = bt.twg_mesh[blockIdx.x];
  if (!bt.mask[mesh]) return;  // uniform over the workgroup
  const int e = blockIdx.x * kHessTile + threadIdx.x;
  elem_hess_tile<T, Model, kHessTile>(m, blockIdx.x, e < bt.tend[mesh], u, dt2, make_spsd, hslot, ht_off, ht_run,
                                      hpart, P);
}

__global__ __launch_bounds__(kBlock) void k_bhess_gather(BatchView bt, int nub, const int* __restrict__ hb_seg,
                                                         const int* __restrict__ hb_pos,
                                                         const double* __restrict__ hpart, long long P,
                                                         const int2* __restrict__ ub_rc,
                                                         const int* __restrict__ ub_pos,
                                                         const unsigned char* __restrict__ dof, int filter_dbc,
                                                         double* __restrict__ val, double* __restrict__ dinv) {
  const int ub = blockIdx.x * kBlock + threadIdx.x;
  if (ub >= nub) return;
  const int2 rc = ub_rc[ub];
  if (!bt.mask[bt.vwg_mesh[rc.x / kBatchRows]]) return;
  double S[9];
  hess_gather_block(ub, hb_seg, hb_pos, hpart, P, S);
  hess_store_block(ub, rc, S, ub_pos, dof, filter_dbc, val, dinv, nullptr);
}

__global__ __launch_bounds__(kBlock) void k_bhess_reduce(BatchView bt, int nub, const int* __restrict__ hseg,
                                                         const double* __restrict__ hcontrib,
                                                         const int2* __restrict__ ub_rc,
                                                         const int* __restrict__ ub_pos,
                                                         const unsigned char* __restrict__ dof, int filter_dbc,
                                                         double* __restrict__ val, double* __restrict__ dinv) {
  const int j0 = blockIdx.x * kBlock, ub = j0 + threadIdx.x;
  const int2 rc = ub < nub ? ub_rc[ub] : make_int2(0, 0);
  const bool active = ub < nub && bt.mask[bt.vwg_mesh[rc.x / kBatchRows]];
  if (!__syncthreads_or(active)) return;  // every mesh of this workgroup has converged
  double S[9];
  staged_run_sum<9>(hseg, j0, min(j0 + kBlock, nub), hcontrib, S);
  if (!active) return;
  hess_store_block(ub, rc, S, ub_pos, dof, filter_dbc, val, dinv, nullptr);
}

// ------------------------------------------------------------------------------------------------------
// adjoint (element part) and conversions between the padded and the compact (caller's) layouts

template <int Model>
__global__ __launch_bounds__(kBlock) void k_belem_adjoint(MeshView m, BatchView bt, const double* __restrict__ u,
                                                          const double* __restrict__ lv, double dt2,
                                                          const int* __restrict__ gperm, double* __restrict__ gcontrib,
                                                          double* __restrict__ g_lam, double* __restrict__ g_mu) {
  const int mesh = bt.twg_mesh[blockIdx.x];
  if (!bt.mask[mesh]) return;
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e < bt.tend[mesh]) elem_adjoint_tet<Model>(m, e, u, lv, dt2, gperm, gcontrib, g_lam, g_mu);
}

// compact vertex v' of mesh m sits at padded vertex kBatchRows·vwg_off[m] + (v' − vcoff[m]); padding gets 0
__global__ __launch_bounds__(kBlock) void k_bpad_vec(BatchView bt, const int* __restrict__ vcoff,
                                                     const int* __restrict__ nvm, const double* __restrict__ compact,
                                                     double* __restrict__ padded) {
  const int mesh = bt.vwg_mesh[blockIdx.x];
  const int v = blockIdx.x * kBlock + threadIdx.x;
  const int local = v - kBatchRows * bt.vwg_off[mesh];
  const bool real = local < nvm[mesh];
  const size_t src = 3 * ((size_t)vcoff[mesh] + local);
#pragma unroll
  for (int a = 0; a < 3; ++a) padded[3 * (size_t)v + a] = real ? compact[src + a] : 0.0;
}

__global__ __launch_bounds__(kBlock) void k_bunpad_vec(BatchView bt, const int* __restrict__ vcoff,
                                                       const int* __restrict__ nvm, const double* __restrict__ padded,
                                                       double* __restrict__ compact) {
  const int mesh = bt.vwg_mesh[blockIdx.x];
  const int v = blockIdx.x * kBlock + threadIdx.x;
  const int local = v - kBatchRows * bt.vwg_off[mesh];
  if (local >= nvm[mesh]) return;
  const size_t dst = 3 * ((size_t)vcoff[mesh] + local);
#pragma unroll
  for (int a = 0; a < 3; ++a) compact[dst + a] = padded[3 * (size_t)v + a];
}

__global__ __launch_bounds__(kBlock) void k_bunpad_tet(BatchView bt, const int* __restrict__ tcoff,
                                                       const double* __restrict__ padded,
                                                       double* __restrict__ compact) {
  const int mesh = bt.twg_mesh[blockIdx.x];
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= bt.tend[mesh]) return;
  compact[(size_t)tcoff[mesh] + (e - kBatchRows * bt.twg_off[mesh])] = padded[e];
}

// ------------------------------------------------------------------------------------------------------
// line-search vector kernels

__global__ __launch_bounds__(kBlock) void k_baxpy_to(BatchView bt, const double* __restrict__ x,
                                                     const double* __restrict__ alpha, const double* __restrict__ d,
                                                     double* __restrict__ out) {
  const int mesh = bt.vwg_mesh[blockIdx.x];
  if (!bt.mask[mesh]) return;
  const double a = alpha[mesh];
#pragma unroll
  for (int j = 0; j < kBVec; ++j) {
    const size_t i = bdof(j);
    out[i] = x[i] + a * d[i];
  }
}

__global__ __launch_bounds__(kBlock) void k_bcopy2(BatchView bt, const double* __restrict__ s1,
                                                   double* __restrict__ d1, const double* __restrict__ s2,
                                                   double* __restrict__ d2) {
  if (!bt.mask[bt.vwg_mesh[blockIdx.x]]) return;
#pragma unroll
  for (int j = 0; j < kBVec; ++j) {
    const size_t i = bdof(j);
    d1[i] = s1[i];
    d2[i] = s2[i];
  }
}

// ------------------------------------------------------------------------------------------------------
// Batched Jacobi-PCG: the single-mesh two-kernel iteration with per-mesh scalars.  The partials of ρ = r·z
// and r·r are per vertex workgroup (pb0 / pb1 alternate with the iteration), those of p·q per 64-row slice;
// each mesh's workgroups sum only the mesh's own partials (a few dozen), in a fixed order.

__global__ __launch_bounds__(kBlock) void k_bpcg_init(BatchView bt, const double* __restrict__ b, double sign,
                                                      const double* __restrict__ dinv, double* __restrict__ x,
                                                      double* __restrict__ r, double* __restrict__ z,
                                                      double* __restrict__ pb) {
  if (!bt.mask[bt.vwg_mesh[blockIdx.x]]) return;
  double v[2] = {0.0, 0.0};
#pragma unroll
  for (int j = 0; j < kBVec; ++j) {
    const size_t i = bdof(j);
    const double bi = sign * b[i];
    const double zi = dinv[i] * bi;
    x[i] = 0.0;
    r[i] = bi;
    z[i] = zi;
    v[0] += bi * zi;
    v[1] += bi * bi;
  }
  store_partials<2>(v, pb);
}

__global__ __launch_bounds__(kBlock) void k_bpcg_init_fin(BatchView bt, const double* __restrict__ pb, double rtol,
                                                          int maxit, PcgScalars* sc, PcgIter* it) {
  const int mesh = blockIdx.x;
  if (mesh == 0 && threadIdx.x == 0) {
    it->ka = 0;
    it->kb = 0;
  }
  PcgScalars* S = sc + mesh;
  if (!bt.mask[mesh]) {
    if (threadIdx.x == 0) {
      S->iters = 0;
      S->breakdown = 0;
      S->done = 1;
    }
    return;
  }
  double t[2];
  const int v0 = bt.vwg_off[mesh];
  sum_partials<2>(pb + 2 * (size_t)v0, bt.vwg_off[mesh + 1] - v0, t);
  if (threadIdx.x == 0) {
    S->rz = t[0];
    S->bb = t[1];
    S->rr = t[1];
    S->tol2 = rtol * rtol * t[1];
    S->iters = 0;
    S->maxit = maxit;
    S->breakdown = 0;
    S->done = (t[1] <= S->tol2 || maxit <= 0 || !(t[1] == t[1])) ? 1 : 0;
  }
}

// kernel A: convergence test on the mesh's ‖r_k‖² (a few dozen partials, so it goes first and a converged
// mesh skips its SpMV), β, s = A z_k on one 64-row slice, p_k, q_k, p_k·q_k.  The grid covers the slices in
// slist[0..nlist) (all slices of the meshes still iterating when the host last compacted the list).  A list entry
// names the slice's mesh and the range of the mesh's partial sums, so the slice's index words, the mesh's scalars
// and both partial sums go out in one round of loads after it.
template <int LPR>
__global__ __launch_bounds__(kSpmvBlock) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_bpcg_spmv(
    BellView A, const int4* __restrict__ slist, int nlist, const double* __restrict__ z,
    double* __restrict__ p, double* __restrict__ q, const double* __restrict__ pb0, const double* __restrict__ pb1,
    PcgScalars* sc, PcgIter* it, double* __restrict__ pa) {
  const int k = it->ka;
  if (blockIdx.x == 0 && threadIdx.x == 0) it->kb = k;
  const int li = logical_block(nlist, A.xcd_remap);
  if (li < 0) return;
  const int4 le = slist[li];  // slice, mesh, the mesh's first vertex workgroup, its vertex workgroups
  const int lb = le.x, v0 = le.z, nw = le.w;
  const int v = lb * kSlice + threadIdx.x;
  FixedRowIndex ix{};
  if (LPR == 0) ix = bell_fixed_index(A, v);
  PcgScalars* S = sc + le.y;
  const int done_before = S->done, maxit = S->maxit;
  const double tol2 = S->tol2;
  double t_cur[2] = {0.0, 0.0}, t_prev[2] = {0.0, 0.0};
  if (k > 0) {
    const double* pb_cur = (k & 1) ? pb1 : pb0;
    const double* pb_prev = (k & 1) ? pb0 : pb1;
    sum_partials_pair<kSpmvBlock>(pb_cur + 2 * (size_t)v0, pb_prev + 2 * (size_t)v0, nw, t_cur, t_prev);
  }
  // the empty asm statement takes the scalars and sums as inputs: the compiler keeps their loads above the test
  asm volatile("" ::"s"(done_before), "s"(maxit), "s"(tol2), "v"(t_cur[0]), "v"(t_cur[1]), "v"(t_prev[0]));
  if (done_before) return;
  double beta = 0.0;
  if (k > 0) {
    const bool nan = !(t_cur[1] == t_cur[1]);
    const bool done = t_cur[1] <= tol2 || k >= maxit || nan;
    if (lb == v0 * kBatchSlices && threadIdx.x == 0) {  // the mesh's first slice reports
      S->iters = k;
      S->rr = t_cur[1];
      S->rz = t_cur[0];
      if (nan) S->breakdown = 2;
      if (done) S->done = 1;
    }
    if (done) return;
    beta = t_cur[0] / t_prev[0];
  }
  double sv[3] = {0.0, 0.0, 0.0};
  // v < nrows always holds here; keeping the row product under a branch, as in the single-mesh kernel, gives it a
  // scheduling region of its own -- without it the compiler spills the fixed-width row at this register budget
  if (v < A.nrows) {
    if (LPR == 0)
      bell_row_fixed(A, v, ix, z, sv);
    else
      row_part<LPR>(A, v, 0, z, sv);
  }
  // the row's z, p and q in one round of loads (no branch around them), as in the single-mesh kernel
  double zo[3], po[3], qo[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const size_t i = 3 * (size_t)v + a;
    zo[a] = z[i];
    po[a] = p[i];
    qo[a] = q[i];
  }
  double acc[1] = {0.0};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
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
  store_partials<1, kSpmvBlock>(acc, pa, lb);
}

// kernel B: α = ρ_k / (p_k·q_k); x += α p, r −= α q, z = D⁻¹ r; partials of (ρ_{k+1}, ‖r_{k+1}‖²).  Workgroup
// i handles vertex workgroup wlist[i].
__global__ __launch_bounds__(kBlock) void k_bpcg_update(const int4* __restrict__ wlist,
                                                        const double* __restrict__ dinv, double* __restrict__ x,
                                                        double* __restrict__ r, double* __restrict__ z,
                                                        const double* __restrict__ p, const double* __restrict__ q,
                                                        const double* __restrict__ pa, double* __restrict__ pb0,
                                                        double* __restrict__ pb1, PcgScalars* sc, PcgIter* it) {
  const int k = it->kb;
  if (blockIdx.x == 0 && threadIdx.x == 0) it->ka = k + 1;
  // the list entry names the mesh and the range of its partial sums: the streaming loads, the done flag and the
  // partial sums then go out in one round (a workgroup of a mesh that finished since the last compaction reads
  // its share of the vectors for nothing)
  const int4 le = wlist[blockIdx.x];  // vertex workgroup, mesh, the mesh's first vertex workgroup, its count
  const int w = le.x, v0 = le.z, nw = le.w;
  PcgScalars* S = sc + le.y;
  double xv[kBVec], rv[kBVec], pv[kBVec], qv[kBVec], dv[kBVec];
#pragma unroll
  for (int j = 0; j < kBVec; ++j) {
    const size_t i = bdof_w(w, j);
    xv[j] = x[i];
    rv[j] = r[i];
    pv[j] = p[i];
    qv[j] = q[i];
    dv[j] = dinv[i];
  }
  const int done_before = S->done;
  const double* pb_cur = (k & 1) ? pb1 : pb0;
  double* pb_next = (k & 1) ? pb0 : pb1;
  double pq[1], rz[2];
  sum_partials_1_2(pa + (size_t)v0 * kBatchSlices, nw * kBatchSlices, pb_cur + 2 * (size_t)v0, nw, pq, rz);
  if (done_before) return;
  if (!(pq[0] > 0.0)) {
    if (w == v0 && threadIdx.x == 0) {
      S->breakdown = 1;
      S->done = 1;
    }
    return;
  }
  const double alpha = rz[0] / pq[0];
  double v[2] = {0.0, 0.0};
#pragma unroll
  for (int j = 0; j < kBVec; ++j) {
    const size_t i = bdof_w(w, j);
    x[i] = xv[j] + alpha * pv[j];
    const double ri = rv[j] - alpha * qv[j];
    const double zi = dv[j] * ri;
    __builtin_nontemporal_store(ri, &r[i]);  // as k_pcg_update
    __builtin_nontemporal_store(zi, &z[i]);
    v[0] += ri * zi;
    v[1] += ri * ri;
  }
  store_partials<2>(v, pb_next, w);
}

}  // namespace

// ------------------------------------------------------------------------------------------------------
// launchers


int launch_belem_grad_tile(const MeshView& m, const BatchView& bt, int ntwg, const double* u, const double* uhat,
                           double dt2, const ushort4* tslot, const int* gt_off, const int* gt_run, double* gpart,
                           double* part_e, bool fp32, hipStream_t s) {
  dispatch_elem(m.model, fp32, [&](auto k) {
    using K = decltype(k);
    k_belem_grad_tile<typename K::type, K::model>
        <<<ntwg, kBlock, 0, s>>>(m, bt, u, uhat, dt2, tslot, gt_off, gt_run, gpart, part_e);
  });
  return ntwg;
}

void launch_bgrad_gather(const BatchView& bt, int nvwg, const int* gv_seg, const int* gv_pos, const double* gpart,
                         const unsigned char* dof, double* g, double* part_g, hipStream_t s) {
  k_bgrad_gather<<<nvwg, kBlock, 0, s>>>(bt, gv_seg, gv_pos, gpart, dof, g, part_g);
}

void launch_bgrad_reduce(const BatchView& bt, int nvwg, const int* gseg, const double* gcontrib,
                         const unsigned char* dof, double* g, double* part_g, hipStream_t s) {
  if (nvwg <= 0) return;
  k_bgrad_reduce<<<nvwg, kBlock, 0, s>>>(bt, gseg, gcontrib, dof, g, part_g);
}

void launch_bfinalize_eval(const BatchView& bt, const double* part_e, const double* part_g, EvalScalars* out,
                           hipStream_t s) {
  if (bt.nmesh <= 0) return;
  k_bfinalize_eval<<<bt.nmesh, kBlock, 0, s>>>(bt, part_e, part_g, out);
}

void launch_bdot(const BatchView& bt, int nvwg, const double* a, const double* b, double* part, EvalScalars* out,
                 hipStream_t s) {
  if (nvwg <= 0) return;
  k_bdot_part<<<nvwg, kBlock, 0, s>>>(bt, a, b, part);
  k_bfinalize_dot<<<bt.nmesh, kBlock, 0, s>>>(bt, part, out);
}

void launch_belem_hess(const MeshView& m, const BatchView& bt, int ntwg, const double* u, double dt2, int make_spsd,
                       const int* hperm, double* hcontrib, bool fp32, hipStream_t s) {
  if (ntwg <= 0) return;
  dispatch_elem(m.model, fp32, [&](auto k) {
    using K = decltype(k);
    k_belem_hess<typename K::type, K::model><<<ntwg, kBlock, 0, s>>>(m, bt, u, dt2, make_spsd, hperm, hcontrib);
  });
}

void launch_belem_hess_tile(const MeshView& m, const BatchView& bt, int ntwg, const double* u, double dt2,
                            int make_spsd, const unsigned* hslot, const int* ht_off, const int* ht_run,
                            double* hpart, long long P, bool fp32, hipStream_t s) {
  if (ntwg <= 0) return;
  dispatch_elem(m.model, fp32, [&](auto k) {
    using K = decltype(k);
    k_belem_hess_tile<typename K::type, K::model>
        <<<ntwg, kHessTile, 0, s>>>(m, bt, u, dt2, make_spsd, hslot, ht_off, ht_run, hpart, P);
  });
}

void launch_bhess_gather(const BatchView& bt, int nub, const int* hb_seg, const int* hb_pos, const double* hpart,
                         long long P, const int2* ub_rc, const int* ub_pos, const unsigned char* dof, int filter_dbc,
                         double* bell_val, double* dinv, hipStream_t s) {
  if (nub <= 0) return;
  k_bhess_gather<<<(nub + kBlock - 1) / kBlock, kBlock, 0, s>>>(bt, nub, hb_seg, hb_pos, hpart, P, ub_rc, ub_pos,
                                                                 dof, filter_dbc, bell_val, dinv);
}

void launch_bhess_reduce(const BatchView& bt, int nub, const int* hseg, const double* hcontrib, const int2* ub_rc,
                         const int* ub_pos, const unsigned char* dof, int filter_dbc, double* bell_val, double* dinv,
                         hipStream_t s) {
  if (nub <= 0) return;
  k_bhess_reduce<<<(nub + kBlock - 1) / kBlock, kBlock, 0, s>>>(bt, nub, hseg, hcontrib, ub_rc, ub_pos, dof,
                                                                 filter_dbc, bell_val, dinv);
}

void launch_belem_adjoint(const MeshView& m, const BatchView& bt, int ntwg, const double* u, const double* lv,
                          double dt2, const int* gperm, double* gcontrib, double* g_lam, double* g_mu, hipStream_t s) {
  if (ntwg <= 0) return;
  if (m.model == kLinearElastic)
    k_belem_adjoint<kLinearElastic><<<ntwg, kBlock, 0, s>>>(m, bt, u, lv, dt2, gperm, gcontrib, g_lam, g_mu);
  else
    k_belem_adjoint<kStableNeoHookean><<<ntwg, kBlock, 0, s>>>(m, bt, u, lv, dt2, gperm, gcontrib, g_lam, g_mu);
}

void launch_bpad_vec(const BatchView& bt, int nvwg, const int* vcoff, const int* nvm, const double* compact,
                     double* padded, hipStream_t s) {
  if (nvwg <= 0) return;
  k_bpad_vec<<<nvwg, kBlock, 0, s>>>(bt, vcoff, nvm, compact, padded);
}

void launch_bunpad_vec(const BatchView& bt, int nvwg, const int* vcoff, const int* nvm, const double* padded,
                       double* compact, hipStream_t s) {
  if (nvwg <= 0) return;
  k_bunpad_vec<<<nvwg, kBlock, 0, s>>>(bt, vcoff, nvm, padded, compact);
}

void launch_bunpad_tet(const BatchView& bt, int ntwg, const int* tcoff, const double* padded, double* compact,
                       hipStream_t s) {
  if (ntwg <= 0) return;
  k_bunpad_tet<<<ntwg, kBlock, 0, s>>>(bt, tcoff, padded, compact);
}

void launch_baxpy_to(const BatchView& bt, int nvwg, const double* x, const double* alpha, const double* d,
                     double* out, hipStream_t s) {
  if (nvwg <= 0) return;
  k_baxpy_to<<<nvwg, kBlock, 0, s>>>(bt, x, alpha, d, out);
}

void launch_bcopy2(const BatchView& bt, int nvwg, const double* src1, double* dst1, const double* src2, double* dst2,
                   hipStream_t s) {
  if (nvwg <= 0) return;
  k_bcopy2<<<nvwg, kBlock, 0, s>>>(bt, src1, dst1, src2, dst2);
}

void launch_bpcg_init(const BatchView& bt, int nvwg, const double* b, double sign, const double* dinv, double* x,
                      double* r, double* z, double rtol, int maxit, PcgScalars* sc, PcgIter* it, double* pb0,
                      hipStream_t s) {
  if (nvwg > 0) k_bpcg_init<<<nvwg, kBlock, 0, s>>>(bt, b, sign, dinv, x, r, z, pb0);
  if (bt.nmesh > 0) k_bpcg_init_fin<<<bt.nmesh, kBlock, 0, s>>>(bt, pb0, rtol, maxit, sc, it);
}

void launch_bpcg_spmv(const BellView& A, const int4* slist, int nlist, const double* z, double* p,
                      double* q, const double* pb0, const double* pb1, PcgScalars* sc, PcgIter* it, double* pa,
                      hipStream_t s) {
  if (nlist <= 0) return;
  const int grid = grid_for(nlist, A.xcd_remap);
  if (A.lanes_per_row == 0)
    k_bpcg_spmv<0><<<grid, kSpmvBlock, 0, s>>>(A, slist, nlist, z, p, q, pb0, pb1, sc, it, pa);
  else
    k_bpcg_spmv<1><<<grid, kSpmvBlock, 0, s>>>(A, slist, nlist, z, p, q, pb0, pb1, sc, it, pa);
}

void launch_bpcg_update(const int4* wlist, int nwl, const double* dinv, double* x, double* r,
                        double* z, const double* p, const double* q, const double* pa, double* pb0, double* pb1,
                        PcgScalars* sc, PcgIter* it, hipStream_t s) {
  if (nwl <= 0) return;
  k_bpcg_update<<<nwl, kBlock, 0, s>>>(wlist, dinv, x, r, z, p, q, pa, pb0, pb1, sc, it);
}

}  // namespace ssim_hip
