// context.h -- state of a libssim_hip.so context (one mesh, or a padded concatenation of meshes for batching),
// shared by context.hip (single-mesh ABI), topology.hip (reset-time setup), solvers.hip (stepper drivers) and
// batch.hip (batched ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ssim_hip.h"
#include "internal.h"


namespace ssim_hip {

extern thread_local std::string g_err;

struct Fail {
  ssim_status_t code;
};

[[noreturn]] void fail(ssim_status_t code, const std::string& msg);

#define HIPCHK(x)                                                                               \
  do {                                                                                          \
    hipError_t e_ = (x);                                                                        \
    if (e_ != hipSuccess) ssim_hip::fail(SSIM_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t count = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    count = 0;
  }
  void alloc(size_t n) {
    if (n == count && p) return;
    release();
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    HIPCHK(hipMalloc(&p, bytes));
    count = n;
  }
  void upload(const T* h, size_t n, hipStream_t s) {
    alloc(n);
    if (n) HIPCHK(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& h, hipStream_t s) { upload(h.data(), h.size(), s); }
};

}  // namespace ssim_hip

struct ssim_ctx {
  int nv = 0, nt = 0, n = 0;
  int device = 0;  // the HIP device current at ssim_create
  double dt = 1e-3, youngs = 1e6, poisson = 0.33, density = 1e3, damping = 0.0;
  double rtol = 1e-2, tol_abs = 1e-2, total_mass = 0.0;
  double rho_reset = 0.0;  // density the mass terms use: captured at reset, as the reference's mass matrix
  bool per_tet = false;
  bool uniform_material = true;  // every tet has (lam0, mu0), bitwise (set by upload_material)
  double lam0 = 0.0, mu0 = 0.0;
  bool has_reset = false;
  ssim_stepper_t stepper = SSIM_STEPPER_NEWTON;
  ssim_material_t model = SSIM_MATERIAL_STABLE_NEOHOOKEAN;
  ssim_solver_params prm{};
  ssim_step_stats stats{};
  hipStream_t stream = nullptr;

  std::vector<double> hX;
  std::vector<int> hcells;
  std::vector<double> hlam, hmu;  // per-tet material when per_tet

  // host topology (block adjacency, rows and columns ascending)
  std::vector<int> brow, bcol, kdiag;
  // blocked-ELL slot of every block of the adjacency: upper blocks (k ≥ kdiag) their upper slot, lower blocks
  // (k < kdiag) their lower-reference slot
  std::vector<int> bslot;
  std::vector<int> hwu, hoffu;  // symmetric blocked ELL: upper widths and slot offsets per slice
  long long ncolslots = 0;      // upper slots (with padding)
  long long nlowslots = 0;      // lower reference slots (with padding)
  int nslices = 0, nub = 0;
  int xcd_remap = 1;
  int pcg_chunk_max = 64;
  int grad_tile = ssim_hip::kBlock;  // tets per tile (= threads per workgroup) of the tile-reduced gradient
  int spmv_lanes = 1;
  bool spmv_lanes_default = true;  // pick the variant from the layout at reset
  bool fixed_layout = false;       // the fixed layout: 7 + 7 off-diagonal slots, 6-value diagonals, 16-bit indices
  // first rows of the independent meshes whose slot directions are chosen separately (a batch's meshes), then nv;
  // empty: one mesh
  std::vector<int> mesh_rows;
  hipEvent_t poll_ev[2] = {nullptr, nullptr};

  // device: mesh and rest data
  ssim_hip::DBuf<double> X, B, V, lam, mu;
  ssim_hip::DBuf<int4> cells;
  // device: vertex arrays
  ssim_hip::DBuf<double> arr[SSIM_ARRAY_COUNT];
  ssim_hip::DBuf<unsigned char> dof;
  ssim_hip::DBuf<double> utry, gtry, dir, r, z, p, q, tmp, dinv, hrest;
  ssim_hip::DBuf<double> ones;  // unit diagonal: the preconditioner of plain CG (pcg_precond = 1)
  // device: assembly maps
  ssim_hip::DBuf<int> gperm, gseg, hperm, hseg, wu, offu, ucol, wl, offl, ub_pos;
  // tile-reduced gradient: tile-local LDS slots of every tet corner, per tile the offsets of its vertices and their
  // run boundaries, per vertex the positions of its (tile, vertex) partials, and the partials themselves
  ssim_hip::DBuf<ushort4> gt_slot;
  ssim_hip::DBuf<int> gt_off, gt_run, gv_seg, gv_pos;
  ssim_hip::DBuf<double> gpart;
  // tile-reduced Hessian: tile-local LDS slots of every (tet, pair) contribution (two 16-bit slots per word, 5 words
  // per tet, word-major), per tile the offsets of its upper blocks and their run boundaries, per upper block the
  // positions of its (tile, block) partials, and the partials, component-major ([9][hparts])
  ssim_hip::DBuf<unsigned> ht_slot;
  ssim_hip::DBuf<int> ht_off, ht_run, hb_seg, hb_pos;
  ssim_hip::DBuf<double> hpart;
  long long hparts = 0;
  int rest_from_x = 1;  // 1: the energy + gradient and tile Hessian kernels form B, V from X (no 80 B/tet loaded)
  int hess_tile = 1;  // 1: tile-reduced Hessian assembly; 0: per-tet contribution slots (k_elem_hess + k_hess_reduce)
  ssim_hip::DBuf<int2> ub_rc, lref;
  ssim_hip::DBuf<uint4> idx;  // fixed layout: two 16-byte index words per row
  ssim_hip::DBuf<double> gcontrib, hcontrib, val, massval;
  ssim_hip::DBuf<double> adj_lam, adj_mu;  // per-tet material gradients when the caller passes no buffer
  // preconditioner factor L of the externally preconditioned NCG: block CSR on the block adjacency
  ssim_hip::DBuf<int> pre_brow, pre_bcol, pre_tpos;
  ssim_hip::DBuf<double> pre_val, pre_a, pre_b, pre_pg;
  // projective-dynamics matrix M + dt²κK (blocked-ELL values and Jacobi) and the two-loop vector of its L-BFGS
  ssim_hip::DBuf<double> pd_val, pd_dinv, pd_q;
  // device: scalars
  ssim_hip::DBuf<ssim_hip::PcgScalars> pcg;
  ssim_hip::DBuf<ssim_hip::PcgIter> pcg_it;
  // captured PCG chunks: `count` iterations of kernels A and B for one output vector x, replayed by one graph
  // launch (the iteration counter is on the device, so every chunk of a given length is the same graph)
  struct PcgGraph {
    int count;
    const double* x;
    const double* val;
    hipGraphExec_t exec;
  };
  std::vector<PcgGraph> pcg_graphs;
  hipStream_t cap_stream = nullptr;  // private stream the chunks are captured on
  int pcg_graph = 0;                 // replay PCG chunks from captured graphs
  int lbfgs_gram = 1;                // two-loop recursion on Gram matrices (diagonal H0, m ≤ kLbfgsSlots)
  ssim_hip::DBuf<ssim_hip::EvalScalars> ev;
  ssim_hip::DBuf<double> partials, lb_S, lb_Y, lb_scal, lb_part;
  size_t pmax = 0;  // partials per slot
  // partial-sum slots: pcg p·q, pcg (r·z, r·r) ping-pong, energy, ‖g‖², dot
  double* part_pa() { return partials.p; }
  double* part_pb(int k) { return partials.p + pmax + 2 * pmax * (size_t)(k & 1); }
  double* part_e() { return partials.p + 5 * pmax; }
  double* part_g() { return partials.p + 6 * pmax; }
  double* part_d() { return partials.p + 7 * pmax; }
  double* pin = nullptr;  // pinned host mirror for scalar read-backs

  // sampled kernel timing: an event pair per kernel, armed at one launch and harvested at the next sync
  struct KTimer {
    hipEvent_t ev[2] = {nullptr, nullptr};
    double sum_ms = 0;
    long long cnt = 0;
    bool pending = false;
  };
  bool timing = false;
  KTimer kt[SSIM_KERNEL_COUNT];

  void drop_graphs() {
    for (auto& g : pcg_graphs) (void)hipGraphExecDestroy(g.exec);
    pcg_graphs.clear();
  }
  ~ssim_ctx() {
    drop_graphs();
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    if (pin) (void)hipHostFree(pin);
    for (auto& e : poll_ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& t : kt)
      for (auto& e : t.ev)
        if (e) (void)hipEventDestroy(e);
  }

  ssim_hip::MeshView mesh() const {
    ssim_hip::MeshView m;
    m.X = X.p;
    m.cells = cells.p;
    m.B = B.p;
    m.V = V.p;
    m.lam = lam.p;
    m.mu = mu.p;
    m.lam0 = lam0;
    m.mu0 = mu0;
    m.uniform = uniform_material ? 1 : 0;
    m.nv = nv;
    m.nt = nt;
    m.rho = rho_reset;
    m.model = static_cast<int>(model);
    m.xcd_remap = xcd_remap;
    m.rest_x = rest_from_x;
    return m;
  }
  ssim_hip::BellView bell(const double* values) const {
    ssim_hip::BellView A;
    A.wu = wu.p;
    A.offu = offu.p;
    A.ucol = ucol.p;
    A.val = values;
    A.wl = wl.p;
    A.offl = offl.p;
    A.lref = lref.p;
    A.idx = idx.p;
    A.nrows = nv;
    A.nslices = nslices;
    A.xcd_remap = xcd_remap;
    A.lanes_per_row = spmv_lanes;
    return A;
  }
  double* a(ssim_array_t w) { return arr[w].p; }
  // value base of upper slot ku of row v (ku = 0: the diagonal block; 6 values in the fixed layout, 9 otherwise)
  size_t valpos(int v, int ku) const {
    const int s = v / ssim_hip::kSlice, lane = v % ssim_hip::kSlice;
    if (fixed_layout)
      return ssim_hip::kBellZeroLead + (size_t)s * ssim_hip::kBellSliceVals +
             (ku ? (size_t)(6 + 9 * (ku - 1)) * ssim_hip::kSlice : 0) + lane;
    return 9 * (size_t)hoffu[s] + (size_t)ku * 9 * ssim_hip::kSlice + lane;
  }
  // offset of entry (b, d) from the value base of a block (a diagonal block of the fixed layout keeps its upper
  // triangle: entry (b, d) and (d, b) share one value)
  static size_t val_entry(bool sym6, int b, int d) {
    if (!sym6) return (size_t)(3 * b + d) * ssim_hip::kSlice;
    const int lo = b < d ? b : d, hi = b < d ? d : b;
    static const int m[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    return (size_t)m[lo][hi] * ssim_hip::kSlice;
  }
};


namespace ssim_hip {
// helpers defined in context.hip, topology.hip and solvers.hip, used across them and by batch.hip
void sync(ssim_ctx* c);
void wait_event(hipEvent_t ev);  // bounded polling, then a blocking wait
void default_params(ssim_solver_params& p);
// marks (mark) or releases vertices as Dirichlet; targets (3 per vertex) may be null for zero targets
void edit_dirichlet(ssim_ctx* c, const int* verts, int count, const double* targets, bool mark);
void harvest_timers(ssim_ctx* c);
// Runs launch(); when timing is on and this kernel's timer is free, brackets it with the timer's events.
template <typename F>
void timed(ssim_ctx* c, ssim_kernel_t k, F&& launch) {
  auto& t = c->kt[k];
  const bool arm = c->timing && !t.pending;
  if (arm) {
    if (!t.ev[0]) {
      HIPCHK(hipEventCreate(&t.ev[0]));
      HIPCHK(hipEventCreate(&t.ev[1]));
    }
    HIPCHK(hipEventRecord(t.ev[0], c->stream));
  }
  launch();
  if (arm) {
    HIPCHK(hipEventRecord(t.ev[1], c->stream));
    t.pending = true;
  }
}

// As timed(), for a launcher that takes the event pair and has the runtime stamp it at the kernel's own start
// and end (hipExtLaunchKernelGGL): the interval then excludes the dispatch gap after the preceding kernel.
template <typename F>
void timed_ext(ssim_ctx* c, ssim_kernel_t k, F&& launch) {
  auto& t = c->kt[k];
  const bool arm = c->timing && !t.pending;
  if (arm && !t.ev[0]) {
    HIPCHK(hipEventCreate(&t.ev[0]));
    HIPCHK(hipEventCreate(&t.ev[1]));
  }
  launch(arm ? t.ev[0] : nullptr, arm ? t.ev[1] : nullptr);
  if (arm) t.pending = true;
}

void do_reset(ssim_ctx* c);
void assemble_hessian(ssim_ctx* c, const double* u, int make_spsd, int filter_dbc, double dt2, double* values);
void assemble_into(ssim_ctx* c, const double* u, double dt2, int make_spsd, int filter_dbc, double* values,
                   double* dinv, double* diag, bool fp32);
void check_material(double youngs, double poisson, double density);  // time_step.hpp:369-374
void require_reset(ssim_ctx* c);
void upload_material(ssim_ctx* c);
void build_grad_tiles(ssim_ctx* c);  // the tile maps of the tile-reduced gradient for grad_tile tets per tile
void build_hess_tiles(ssim_ctx* c, const std::vector<int>& hub);  // the tile maps of the tile-reduced Hessian
void ensure_hcontrib(ssim_ctx* c);  // the per-tet contribution slots (projective dynamics, hess_tile = 0)

// stepper drivers (solvers.hip)
struct EvalOut {
  double E, gg;
};
EvalOut eval_energy_grad(ssim_ctx* c, const double* u, double* g);
// work buffers of one PCG solve: residual r, preconditioned residual z, direction p, q = A p, the partial sums of
// kernel A (pa) and of the residual ping-pong (pb0, pb1), and the device-side scalars and iteration counter
struct PcgWork {
  double *r, *z, *p, *q, *pa, *pb0, *pb1;
  PcgScalars* sc;
  PcgIter* it;
};
PcgWork ctx_pcg_work(ssim_ctx* c);
void create_poll_events(ssim_ctx* c);  // the two events pcg_solve polls (created once, by reset)
void ensure_unit_diagonal(ssim_ctx* c);  // the all-ones diagonal of plain CG (pcg_precond = 1)
void noalloc_scratch_init();             // noalloc.hip: the current device's address of the no-alloc scratch
int pcg_solve(ssim_ctx* c, const double* b, double sign, double* x, double rtol, int maxit,
              const double* val = nullptr, const double* dinv = nullptr, const PcgWork* work = nullptr);
void newton_solve(ssim_ctx* c);
void lbfgs_solve(ssim_ctx* c, bool pd, bool h0 = false);
void ncg_solve(ssim_ctx* c, bool precond);
void forward_euler_solve(ssim_ctx* c);
void set_precond(ssim_ctx* c, const double* values);
void build_pd(ssim_ctx* c);
template <typename F>
ssim_status_t run(F&& f) {
  try {
    f();
    return SSIM_OK;
  } catch (const Fail& e) {
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host out of memory";
    return SSIM_ERR_HIP;
  } catch (const std::exception& e) {
    g_err = e.what();
    return SSIM_ERR_STATE;
  }
}
}  // namespace ssim_hip
