// internal.h -- device data views and kernel launchers shared by the HIP sources of libssim_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ssim_hip {

constexpr int kBlock = 256;      // threads per workgroup for every kernel (4 wave64)
constexpr int kSlice = 64;       // rows per blocked-ELL slice (one wavefront)
constexpr double kSpdFloor = 1e-3;  // eigenvalue floor of the SPD projection (time_step.hpp:52)
constexpr int kMinTile = 64;  // the smallest tile of the tile-reduced gradient (workgroup partials are sized for it)
constexpr int kFixedWU = 8;  // uniform upper / lower slot counts of the fixed-width SpMV layout
constexpr int kFixedWL = 7;

// Per-tet rest data in HBM.  cells are int4 (one 16-byte load per tet); B = Dm⁻¹ is structure-of-arrays
// B[k·nt + e] for k = 0..8 so that a wavefront's loads of one component are contiguous.
struct MeshView {
  const double* X;      // [nv*3] rest positions
  const int4* cells;    // [nt]
  const double* B;      // [9][nt]
  const double* V;      // [nt]  rest volumes
  const double* lam;    // [nt]
  const double* mu;     // [nt]
  double lam0, mu0;     // the material when every tet has the same (λ, μ); the arrays are then not read
  int uniform;          // 1: use (lam0, mu0)
  int nv, nt;
  double rho;
  int model = 0;      // MaterialModel (elem_math.h)
  int xcd_remap = 0;  // 1: the element kernels give each of the 8 XCDs a contiguous range of tets
  int rest_x = 0;     // 1: the energy + gradient and tile Hessian kernels form B and V from the gathered rest
                      // positions (the same arithmetic as the rest-shape kernel) instead of loading them
};

// Symmetric blocked ELLPACK of 3x3 blocks.  Rows (vertices) are grouped in slices of 64 lanes.  Two layouts:
//
// General (rows of any width; read by the looped row products):
// Upper part (columns w ≥ v, diagonal first): slice s has wu[s] slots,
//   ucol[offu[s] + k*64 + lane] = column vertex,
//   val[9*offu[s] + (k*9 + 3b + d)*64 + lane] = entry (b, d) of block (v, w).
// Lower part (columns w < v) is not stored: slot k of slice s holds lref[offl[s] + k*64 + lane] =
//   (w, base) with val[base + (3b + d)*64] = entry (b, d) of the stored block (w, v), used transposed.
// Padding: upper slots hold column v with zero values; lower references hold (v, base of a zero block that
// follows the last slice), so the SpMV loops need no branch.
//
// Fixed (every row has at most kFixedWU − 1 = 7 upper and kFixedWL = 7 lower off-diagonal blocks, and every
// column is within 65,535 vertices of its row; read by the unrolled row product):
//   val[kBellZeroLead + s*kBellSliceVals + m*64 + lane], m = 0..5: the diagonal block's upper triangle
//     (00, 01, 02, 11, 12, 22) -- the stored Hessian is symmetric, so the diagonal keeps 6 of its 9 values;
//   val[kBellZeroLead + s*kBellSliceVals + (6 + 9(k−1) + 3b + d)*64 + lane]: entry (b, d) of upper slot
//     k = 1..7;
//   val[0 .. kBellZeroLead): a zero block, read (lane-strided) by padded lower references;
//   idx[(2s + h)*64 + lane], h = 0, 1: the row's 16-bit column offsets and lower references, 32 B per row:
//     h = 0: (u1 | u2<<16, u3 | u4<<16, u5 | u6<<16, u7 | l0<<16),
//     h = 1: (l1 | l2<<16, l3 | l4<<16, l5 | l6<<16, slot word),
//   with uk = column − v of upper slot k (0 for padding: column v, zero values), lj = v − w of lower slot j,
//   and bits 3j..3j+2 of the slot word the upper slot (1..7) of row w that stores block (w, v) (0 for
//   padding).  Against the general layout this moves 80 B less per row (diagonal 24 B, indices 56 B).
constexpr int kBellZeroLead = 9 * kSlice;
constexpr int kBellSliceVals = (6 + 9 * (kFixedWU - 1)) * kSlice;
constexpr int kBellMaxOffset = 65535;

struct BellView {
  const int* wu;
  const int* offu;
  const int* ucol;
  const double* val;
  const int* wl;
  const int* offl;
  const int2* lref;
  const uint4* idx;  // fixed layout only
  int nrows, nslices;
  int xcd_remap;      // 1: give each of the 8 XCDs a contiguous range of rows
  int lanes_per_row;  // 0 (fixed layout), 1 or 4 (general layout)
};

// Scalars of one PCG solve, resident in device memory.
struct PcgScalars {
  double rz, pq, alpha, beta, rr, bb, tol2, pad;
  double rz_hist[2];  // ρ_k in slot k % 2, written by kernel A of iteration k (ρ_0 by the initialisation)
  int done, iters, maxit, breakdown;  // read back together by the host
  // Iteration at which the single-mesh PCG stopped (INT_MAX while it runs, −1 if it stops before iterating).
  // Kernel A returns at entry only when its iteration is past stop_k, i.e. when an earlier launch stopped the
  // solve; a stop published during the same launch (by its workgroup 0) never makes a workgroup skip its rows.
  int stop_k, pad_i;
};

// Reduction results of element / vector kernels.
struct EvalScalars {
  double energy, gg, gd, dot;
};

// ---- batches of meshes ----
// A batch concatenates its meshes into one mesh whose per-mesh vertex and tet ranges are padded to whole
// workgroups of kBatchRows: every vector / row-slice / element workgroup then belongs to exactly one mesh, so
// per-workgroup partial sums are per-mesh partial sums, and a mesh that sits out a launch (converged, or done
// with its line search) skips it at workgroup granularity.  Padding vertices are Dirichlet-fixed and in no tet;
// padding tets are absent (first vertex index −1).
constexpr int kBatchRows = 256;         // vertex and tet granule of a mesh (= kBlock)
constexpr int kBatchSlices = kBatchRows / kSlice;  // blocked-ELL slices per vertex workgroup
constexpr int kHessTile = kBatchRows;  // tets per tile of the tile-reduced Hessian (a batch tile is in one mesh)
struct BatchView {
  int nmesh;
  const int* vwg_mesh;  // [vertex workgroups] mesh of each block of kBatchRows vertices
  const int* twg_mesh;  // [tet workgroups] mesh of each block of kBatchRows tets
  const int* vwg_off;   // [nmesh+1] first vertex workgroup of each mesh
  const int* twg_off;   // [nmesh+1] first tet workgroup of each mesh
  const int* tend;      // [nmesh] one past the last real tet of each mesh
  const int* mask;      // [nmesh] nonzero: the mesh takes part in the launch
};
// Iteration counter of a batched PCG, kept on the device so that every iteration's launches are identical
// (replayable from a captured graph): kernel A reads ka and publishes kb = ka, kernel B reads kb and
// publishes ka = kb + 1; neither kernel writes the field it reads.
struct PcgIter {
  int ka, kb;
};

// ---- launchers (kernels.hip) ----
void launch_rest_shape(const double* X, const int4* cells, int nt, double* B, double* V, hipStream_t s);
// Reductions: a kernel stores one partial per workgroup; launch_* return the number of partials written.
// Launchers that take an event pair (e0, e1) have the runtime stamp it at the kernel's own start and end.
// energy + gradient at deformation u (predictor uhat), reduced per tile of kBlock tets in LDS: one partial per
// (tile, vertex of the tile); energy partials to part_e
int launch_elem_grad_tile(const MeshView& m, int tile, const double* u, const double* uhat, double dt2,
                          const ushort4* tslot, const int* gt_off, const int* gt_run, double* gpart, double* part_e,
                          bool fp32, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// g[v] = Σ of v's tile partials (in tile order), zeroed on Dirichlet dofs; ‖g‖² partials to part_g
int launch_grad_gather(int nv, const int* gv_seg, const int* gv_pos, const double* gpart, const unsigned char* dof,
                       double* g, double* part_g, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// g[v] = Σ per-corner contributions of v (vertex-sorted slots gperm, runs gseg; the adjoint's mass term), zeroed
// on Dirichlet dofs; ‖g‖² partials to part_g
int launch_grad_reduce(int nv, const int* gseg, const double* gcontrib, const unsigned char* dof, double* g,
                       double* part_g, hipStream_t s);
// one workgroup: out[0..ka) = Σ pa (width ka ∈ {1,2}); if pb: out[ka] = Σ pb (width 1)
void launch_finalize(const double* pa, int na, int ka, const double* pb, int nb, double* out, hipStream_t s);
// Hessian contributions (10 oriented 3x3 blocks per tet, dt²V-scaled, consistent mass on the diagonal blocks)
void launch_elem_hess(const MeshView& m, const double* u, double dt2, int make_spsd, const int* hperm,
                      double* hcontrib, bool fp32, hipStream_t s);
void launch_hess_reduce(int nub, const int* hseg, const double* hcontrib, const int2* ub_rc, const int* ub_pos,
                        const unsigned char* dof, int filter_dbc, double* bell_val, double* dinv, double* diag,
                        hipStream_t s);
// tile-reduced Hessian assembly: one partial per (tile of kHessTile tets, upper block) and component into
// hpart[comp·P + position], then per upper block the sum of its partials in tile order, the Dirichlet identity
// filter, the blocked-ELL store and the Jacobi inverse diagonal (and the diagonal when diag is not null)
void launch_elem_hess_tile(const MeshView& m, const double* u, double dt2, int make_spsd, const unsigned* hslot,
                           const int* ht_off, const int* ht_run, double* hpart, long long P, bool fp32,
                           hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void launch_hess_gather(int nub, const int* hb_seg, const int* hb_pos, const double* hpart, long long P,
                        const int2* ub_rc, const int* ub_pos, const unsigned char* dof, int filter_dbc,
                        double* bell_val, double* dinv, double* diag, hipStream_t s, hipEvent_t e0 = nullptr,
                        hipEvent_t e1 = nullptr);
void launch_elem_eval(const MeshView& m, const double* u, int make_spsd, double* psi, double* P, double* H,
                      double* K, bool fp32, hipStream_t s);

void launch_spmv(const BellView& A, const double* x, double* y, hipStream_t s);
void launch_dot(int n, const double* a, const double* b, double* part, double* out, hipStream_t s);
void launch_prepare(int n, const double* u, const double* vel, const double* accel, const unsigned char* dof,
                    const double* dbc, double dt, double* uhat, double* unext, hipStream_t s);
void launch_step_next(int n, const double* unext, double* u, double* uprev, double* vel, double dt,
                      double damping, hipStream_t s);
void launch_restore(int n, const unsigned char* dof, const double* dbc, double* u, double* uprev, double* vel,
                    double dt, hipStream_t s);
void launch_axpy_to(int n, const double* x, double alpha, const double* d, double* out, hipStream_t s);

// PCG on H x = sign·b with Jacobi preconditioner dinv.  init writes (ρ0, ‖r0‖²) partials to pb0, resets the
// device iteration counter and returns the partial count.  Iteration k (k from the counter): kernel A (s = A z_k;
// convergence test on pb[k%2]; p_k = z_k + β p_{k−1}, q_k = s + β q_{k−1} in place; partials of p_k·q_k to pa),
// kernel B (reads pa and pb[k%2], writes pb[(k+1)%2]).  Every iteration's launches are identical.
int launch_pcg_init(int n, const double* b, double sign, const double* dinv, double* x, double* r, double* z,
                    double rtol, int maxit, PcgScalars* sc, PcgIter* it, double* pb0, hipStream_t s);
int launch_pcg_spmv(const BellView& A, const double* z, double* p, double* q, const double* pb0, const double* pb1,
                    int nb, PcgIter* it, PcgScalars* sc, double* pa, double* x, hipStream_t s,
                    hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void launch_pcg_update(int n, const double* dinv, double* r, double* z, const double* q,
                       const double* pa, int na, double* pb0, double* pb1, PcgIter* it, PcgScalars* sc,
                       hipStream_t s,
                       hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
int pcg_vector_blocks(int n);
int pcg_spmv_grid(const BellView& A);  // workgroups (= p·q partials) of one kernel-A launch

// adjoint of one step: mass contributions of the adjoint vector lv (reduce with launch_grad_reduce to get M·lv,
// zero on Dirichlet dofs) and per-tet ∂L/∂λ, ∂L/∂μ at deformation u
void launch_elem_adjoint(const MeshView& m, const double* u, const double* lv, double dt2, const int* gperm,
                         double* gcontrib, double* g_lam, double* g_mu, hipStream_t s);
// out = in with Dirichlet dofs zeroed
void launch_mask_dbc(int n, const unsigned char* dof, const double* in, double* out, hipStream_t s);

// nonlinear CG: out[0..5) = (dᵀy, gtᵀy, yᵀy, gtᵀd, dᵀd), y = gt − g (part holds vector-block partials of width 5)
void launch_ncg_dots(int n, const double* gt, const double* g, const double* d, const double* a, const double* b,
                     double* part, double* out, hipStream_t s);
void launch_material_eval(int dim, int model, long long n, const double* F, const double* lam, const double* mu,
                          int make_spsd, double* psi, double* P, double* H, hipStream_t s);
// outputs of the P1 triangle evaluation, per triangle (each may be null): area, Dm⁻¹ (row-major 2x2), F
// (row-major), A·Ψ, vertex gradient [3][2], stiffness [6][6] (dof 2a+c), consistent mass [3][3], Laplacian [3][3]
struct TriOut {
  double *area, *dminv, *F, *energy, *grad, *stiff, *mass, *laplacian;
};
void launch_tri_eval(int model, int nv, int nt, const double* X, const int* cells, const double* u,
                     const double* lam, const double* mu, double rho, int make_spsd, const TriOut& o, hipStream_t s);
// L-BFGS in Gram form: at most kLbfgsSlots pairs; kLbfgsDots dot products per iteration (see kernels.hip)
constexpr int kLbfgsSlots = 8;
constexpr int kLbfgsDots = 6 * kLbfgsSlots + 1;
struct LbfgsCoef {
  double a[kLbfgsSlots], c[kLbfgsSlots];
  unsigned mask;
};
// k ≥ 0: the pair of slot k is formed from (u − u_old, g − g_old) and stored there by the same pass
void launch_lbfgs_gram(int n, double* S, double* Y, const double* g, const double* h, int k, const double* u,
                       const double* u_old, const double* g_old, double* part, double* out, hipStream_t s);
int lbfgs_gram_blocks(int n);  // partial rows (of kLbfgsDots doubles) that launch_lbfgs_gram writes
// also writes the first trial point ut = u + alpha0·r
void launch_lbfgs_dir(int n, const double* S, const double* Y, const double* g, const double* h, const LbfgsCoef& cf,
                      double* r, const double* u, double alpha0, double* ut, hipStream_t s);
void launch_lbfgs_shift(int n, const double* ut, const double* gt, double* u, double* g, double* sv, double* yv,
                        hipStream_t s);
void launch_elem_pd(const MeshView& m, double c, const int* hperm, double* hcontrib, hipStream_t s);
void launch_bsr_spmv(int nv, const int* brow, const int* bcol, const int* tpos, const double* val, const double* x,
                     double* y, bool trans, hipStream_t s);
// y = a·x + b·y (y is not read when b = 0)
void launch_axpby(int n, double a, const double* x, double b, double* y, hipStream_t s);

// L-BFGS helpers
void launch_scale_div(int n, const double* q, const double* diag, double* out, hipStream_t s);
// y += (c1·(*p1) − c2·(*p2))·x, coefficients read from device memory (p2 may be null)
void launch_axpy_coef(int n, double c1, const double* p1, double c2, const double* p2, const double* x, double* y,
                      hipStream_t s);
void launch_sub(int n, const double* a, const double* b, double* out, hipStream_t s);

int max_partials(int n);  // upper bound on partials any kernel writes for n items

// ---- batched launchers (batch_kernels.hip); nvwg / ntwg = vertex / tet workgroups of the batch ----
// tile-reduced batched evaluation: one partial per (tile, vertex); the gather adds them per vertex
int launch_belem_grad_tile(const MeshView& m, const BatchView& bt, int ntwg, const double* u, const double* uhat,
                           double dt2, const ushort4* tslot, const int* gt_off, const int* gt_run, double* gpart,
                           double* part_e, bool fp32, hipStream_t s);
void launch_bgrad_gather(const BatchView& bt, int nvwg, const int* gv_seg, const int* gv_pos, const double* gpart,
                         const unsigned char* dof, double* g, double* part_g, hipStream_t s);
void launch_bgrad_reduce(const BatchView& bt, int nvwg, const int* gseg, const double* gcontrib,
                         const unsigned char* dof, double* g, double* part_g, hipStream_t s);
// per mesh (masked): out[m].energy = Σ part_e over its tet workgroups, out[m].gg = Σ part_g over its vertex
// workgroups (part_g may be null)
void launch_bfinalize_eval(const BatchView& bt, const double* part_e, const double* part_g, EvalScalars* out,
                           hipStream_t s);
// per mesh (masked): out[m].gd = a·b over the mesh's dofs
void launch_bdot(const BatchView& bt, int nvwg, const double* a, const double* b, double* part, EvalScalars* out,
                 hipStream_t s);
void launch_belem_hess(const MeshView& m, const BatchView& bt, int ntwg, const double* u, double dt2, int make_spsd,
                       const int* hperm, double* hcontrib, bool fp32, hipStream_t s);
void launch_bhess_reduce(const BatchView& bt, int nub, const int* hseg, const double* hcontrib, const int2* ub_rc,
                         const int* ub_pos, const unsigned char* dof, int filter_dbc, double* bell_val, double* dinv,
                         hipStream_t s);
// tile-reduced batched Hessian assembly (masked meshes only; tet workgroup b is tile b)
void launch_belem_hess_tile(const MeshView& m, const BatchView& bt, int ntwg, const double* u, double dt2,
                            int make_spsd, const unsigned* hslot, const int* ht_off, const int* ht_run,
                            double* hpart, long long P, bool fp32, hipStream_t s);
void launch_bhess_gather(const BatchView& bt, int nub, const int* hb_seg, const int* hb_pos, const double* hpart,
                         long long P, const int2* ub_rc, const int* ub_pos, const unsigned char* dof, int filter_dbc,
                         double* bell_val, double* dinv, hipStream_t s);
void launch_belem_adjoint(const MeshView& m, const BatchView& bt, int ntwg, const double* u, const double* lv,
                          double dt2, const int* gperm, double* gcontrib, double* g_lam, double* g_mu, hipStream_t s);
// compact (caller's, unpadded, meshes in order) ↔ padded layouts; vcoff / tcoff = compact vertex / tet offset of
// each mesh, nvm = its vertex count; padding entries of padded vectors are written as 0
void launch_bpad_vec(const BatchView& bt, int nvwg, const int* vcoff, const int* nvm, const double* compact,
                     double* padded, hipStream_t s);
void launch_bunpad_vec(const BatchView& bt, int nvwg, const int* vcoff, const int* nvm, const double* padded,
                       double* compact, hipStream_t s);
void launch_bunpad_tet(const BatchView& bt, int ntwg, const int* tcoff, const double* padded, double* compact,
                       hipStream_t s);
// out = x + alpha[m]·d on the dofs of masked meshes
void launch_baxpy_to(const BatchView& bt, int nvwg, const double* x, const double* alpha, const double* d,
                     double* out, hipStream_t s);
// dst1 = src1, dst2 = src2 on the dofs of masked meshes
void launch_bcopy2(const BatchView& bt, int nvwg, const double* src1, double* dst1, const double* src2, double* dst2,
                   hipStream_t s);
// batched PCG: per-mesh scalars sc[m]; unmasked meshes are marked done at init
void launch_bpcg_init(const BatchView& bt, int nvwg, const double* b, double sign, const double* dinv, double* x,
                      double* r, double* z, double rtol, int maxit, PcgScalars* sc, PcgIter* it, double* pb0,
                      hipStream_t s);
// iteration kernels over compacted grids: slist[0..nlist) = the 64-row slices, wlist[0..nwl) = the vertex
// workgroups of the meshes still iterating (a superset is correct: a done mesh's workgroups return at entry).
// Each entry is {slice or vertex workgroup, mesh, the mesh's first vertex workgroup, its number of vertex
// workgroups}, so a workgroup has its mesh and the range of the mesh's partial sums in its first load.
void launch_bpcg_spmv(const BellView& A, const int4* slist, int nlist, const double* z, double* p,
                      double* q, const double* pb0, const double* pb1, PcgScalars* sc, PcgIter* it, double* pa,
                      hipStream_t s);
void launch_bpcg_update(const int4* wlist, int nwl, const double* dinv, double* x, double* r,
                        double* z, const double* p, const double* q, const double* pa, double* pb0, double* pb1,
                        PcgScalars* sc, PcgIter* it, hipStream_t s);

}  // namespace ssim_hip
