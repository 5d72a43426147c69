// context.hip -- host side of libssim_hip.so: context state, setup of the assembly maps and the blocked-ELL
// matrix layout, the Newton / L-BFGS drivers and the extern "C" ABI declared in include/ssim_hip.h.
//
// The host drives control flow only (Newton iterations, line-search decisions, PCG chunking); every
// floating-point operation of a step runs in the HIP kernels of kernels.hip.  PCG scalars stay on the device
// and the host polls a done-flag once per chunk of iterations.
#include "context.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace ssim_hip;

namespace ssim_hip {
thread_local std::string g_err = "";
[[noreturn]] void fail(ssim_status_t code, const std::string& msg) {
  g_err = msg;
  throw Fail{code};
}
}  // namespace ssim_hip

void ssim_hip::default_params(ssim_solver_params& p) {
  p.max_newton = 10;
  p.pcg_rtol = 1e-8;
  p.pcg_maxit = 10000;
  p.ls_c = 0.3;
  p.ls_beta = 0.5;
  p.ls_max = 20;
  p.lbfgs_m = 8;
  p.lbfgs_maxit = 1000;
  p.lbfgs_c = 1e-4;
  p.lbfgs_ls_max = 30;
  p.fp32_elements = 0;
  p.newton_polish = 0;
}

namespace {

void check_material(double youngs, double poisson, double density) {
  // time_step.hpp:369-374
  if (!(youngs >= 0)) fail(SSIM_ERR_INVALID_ARGUMENT, "Young's modulus must be positive, got " + std::to_string(youngs));
  if (!(poisson > 0 && poisson < 0.5))
    fail(SSIM_ERR_INVALID_ARGUMENT, "Poisson's ratio must be in (0, 0.5), got " + std::to_string(poisson));
  if (!(density >= 0)) fail(SSIM_ERR_INVALID_ARGUMENT, "Density must be positive, got " + std::to_string(density));
}

}  // namespace

void ssim_hip::harvest_timers(ssim_ctx* c) {
  for (auto& t : c->kt) {
    if (!t.pending) continue;
    if (hipEventQuery(t.ev[1]) != hipSuccess) continue;  // not finished yet: harvested at a later sync
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, t.ev[0], t.ev[1]));
    t.sum_ms += ms;
    t.cnt++;
    t.pending = false;
  }
}

void ssim_hip::sync(ssim_ctx* c) {
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->timing) harvest_timers(c);
}

namespace {

// Blocked-ELL slots of every block.  By default the upper blocks of a row take slots 0, 1, ... in column order
// (diagonal first) and the lower references likewise.  In the fixed 8/7 layout the slots are aligned by
// direction instead: slot k ≥ 1 holds, where it can, the block at the column offset w − v that sits at sorted
// position k in most full rows (on a structured mesh every row then keeps each neighbour direction in the same
// slot), and the lower reference to the block stored in slot k of the neighbour's row goes to lower slot 7 − k.
// The fixed-width SpMV pairs upper slot k with lower slot 7 − k, so that for most rows of a slice the lower
// reads and the neighbouring slice's own reads of the same lines are issued at the same point of the program
// and meet in L2.  Slots only move blocks; every block is still read and summed once per row.
void assign_slots(ssim_ctx* c, const std::vector<int>& len) {
  const int nv = c->nv;
  c->bslot.assign(std::max(c->brow[nv], 1), 0);
  for (int v = 0; v < nv; ++v) {
    const int kd = c->kdiag[v];
    for (int k = 0; k < len[v]; ++k) c->bslot[c->brow[v] + k] = k < kd ? k : k - kd;
  }
  if (!c->fixed_layout) return;
  static_assert(kFixedWL == kFixedWU - 1, "the slot pairing needs one lower slot per off-diagonal upper slot");
  // modal column offset of each sorted upper position over the rows with all kFixedWU upper blocks
  std::vector<std::vector<int>> offs(kFixedWU);
  for (int v = 0; v < nv; ++v)
    if (len[v] - c->kdiag[v] == kFixedWU)
      for (int q = 1; q < kFixedWU; ++q) offs[q].push_back(c->bcol[c->brow[v] + c->kdiag[v] + q] - v);
  std::vector<long long> modal(kFixedWU, -1);  // −1: no canonical offset (offsets are positive)
  for (int q = 1; q < kFixedWU; ++q) {
    auto& o = offs[q];
    if (o.empty()) continue;
    std::sort(o.begin(), o.end());
    int best = o[0], run = 0, best_run = 0;
    for (size_t i = 0; i < o.size(); ++i) {
      run = (i > 0 && o[i] == o[i - 1]) ? run + 1 : 1;
      if (run > best_run) {
        best_run = run;
        best = o[i];
      }
    }
    modal[q] = best;
    o.clear();
    o.shrink_to_fit();
  }
  // upper slots: the diagonal in slot 0, canonical directions in their slot, the rest in the free slots
  for (int v = 0; v < nv; ++v) {
    const int kd = c->kdiag[v];
    bool used[kFixedWU] = {true};
    int rest[kFixedWU], nrest = 0;
    c->bslot[c->brow[v] + kd] = 0;
    for (int k = kd + 1; k < len[v]; ++k) {
      const long long off = c->bcol[c->brow[v] + k] - v;
      int q = 1;
      while (q < kFixedWU && !(modal[q] == off && !used[q])) ++q;
      if (q < kFixedWU) {
        c->bslot[c->brow[v] + k] = q;
        used[q] = true;
      } else {
        rest[nrest++] = k;
      }
    }
    for (int i = 0, q = 1; i < nrest; ++i) {
      while (used[q]) ++q;
      c->bslot[c->brow[v] + rest[i]] = q;
      used[q] = true;
    }
  }
  // lower references: the block stored in slot k of row w goes to lower slot kFixedWU − 1 − k where free
  auto upper_slot_of = [&](int w, int v) -> int {
    const int* b = c->bcol.data() + c->brow[w];
    const int* e = c->bcol.data() + c->brow[w + 1];
    const int* it = std::lower_bound(b, e, v);
    if (it == e || *it != v) fail(SSIM_ERR_STATE, "internal: missing block in sparsity pattern");
    return c->bslot[c->brow[w] + (int)(it - b)];
  };
  for (int v = 0; v < nv; ++v) {
    const int kd = c->kdiag[v];
    bool used[kFixedWL] = {};
    int rest[kFixedWL], nrest = 0;
    for (int k = 0; k < kd; ++k) {
      const int j = kFixedWU - 1 - upper_slot_of(c->bcol[c->brow[v] + k], v);
      if (j >= 0 && j < kFixedWL && !used[j]) {
        c->bslot[c->brow[v] + k] = j;
        used[j] = true;
      } else {
        rest[nrest++] = k;
      }
    }
    for (int i = 0, j = 0; i < nrest; ++i) {
      while (used[j]) ++j;
      c->bslot[c->brow[v] + rest[i]] = j;
      used[j] = true;
    }
  }
}

// Sparsity, blocked-ELL layout and the two assembly permutations.
void build_topology(ssim_ctx* c) {
  const int nv = c->nv, nt = c->nt;
  const int* cells = c->hcells.data();
  // block adjacency: neighbours (self included, also for a vertex in no tet) of every vertex, ascending.
  // Tets with a negative first index are absent (padding of a batched layout) and take no part.
  std::vector<int> cnt(nv + 1, 1);
  cnt[0] = 0;
  for (int e = 0; e < nt; ++e)
    if (cells[4 * e] >= 0)
      for (int i = 0; i < 4; ++i) cnt[cells[4 * e + i] + 1] += 4;
  for (int v = 0; v < nv; ++v) cnt[v + 1] += cnt[v];
  std::vector<int> cand(std::max(cnt[nv], 1));
  std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
  for (int v = 0; v < nv; ++v) cand[fillp[v]++] = v;
  for (int e = 0; e < nt; ++e)
    if (cells[4 * e] >= 0)
      for (int i = 0; i < 4; ++i) {
        const int v = cells[4 * e + i];
        for (int j = 0; j < 4; ++j) cand[fillp[v]++] = cells[4 * e + j];
      }
  c->brow.assign(nv + 1, 0);
  std::vector<int> len(nv, 0);
  for (int v = 0; v < nv; ++v) {
    int* b = cand.data() + cnt[v];
    int* e = cand.data() + cnt[v + 1];
    std::sort(b, e);
    len[v] = (int)(std::unique(b, e) - b);
    c->brow[v + 1] = c->brow[v] + len[v];
  }
  c->bcol.resize(std::max(c->brow[nv], 1));
  for (int v = 0; v < nv; ++v) std::copy(cand.begin() + cnt[v], cand.begin() + cnt[v] + len[v], c->bcol.begin() + c->brow[v]);
  cand.clear();
  cand.shrink_to_fit();
  auto slot = [&](int v, int w) -> int {  // index of w in the sorted neighbour list of v
    const int* b = c->bcol.data() + c->brow[v];
    const int* e = c->bcol.data() + c->brow[v + 1];
    const int* it = std::lower_bound(b, e, w);
    if (it == e || *it != w) fail(SSIM_ERR_STATE, "internal: missing block in sparsity pattern");
    return (int)(it - b);
  };
  c->kdiag.resize(nv);
  for (int v = 0; v < nv; ++v) c->kdiag[v] = slot(v, v);

  // symmetric blocked ELL: slices of 64 rows; the upper part (w ≥ v, diagonal first) is stored, the lower
  // part refers to the stored transposes
  c->nslices = (nv + kSlice - 1) / kSlice;
  const int ns = std::max(c->nslices, 1);
  c->hwu.assign(ns, 0);
  c->hoffu.assign(ns + 1, 0);
  std::vector<int> hwl(ns, 0), hoffl(ns + 1, 0);
  int max_wu = 0, max_wl = 0;
  for (int s = 0; s < c->nslices; ++s) {
    for (int l = 0; l < kSlice && s * kSlice + l < nv; ++l) {
      const int v = s * kSlice + l;
      c->hwu[s] = std::max(c->hwu[s], len[v] - c->kdiag[v]);
      hwl[s] = std::max(hwl[s], c->kdiag[v]);
    }
    max_wu = std::max(max_wu, c->hwu[s]);
    max_wl = std::max(max_wl, hwl[s]);
  }
  // rows narrow enough for the fixed-width kernel: pad every slice to the same widths
  c->fixed_layout = max_wu <= kFixedWU && max_wl <= kFixedWL;
  if (c->fixed_layout)
    for (int s = 0; s < c->nslices; ++s) {
      c->hwu[s] = kFixedWU;
      hwl[s] = kFixedWL;
    }
  if (c->spmv_lanes == 0 && !c->fixed_layout) c->spmv_lanes = 1;
  if (c->spmv_lanes_default) c->spmv_lanes = c->fixed_layout ? 0 : 1;
  assign_slots(c, len);
  long long offu = 0, offl = 0;
  for (int s = 0; s < c->nslices; ++s) {
    c->hoffu[s] = (int)offu;
    hoffl[s] = (int)offl;
    offu += (long long)c->hwu[s] * kSlice;
    offl += (long long)hwl[s] * kSlice;
  }
  c->ncolslots = offu;
  c->nlowslots = offl;

  if (9 * offu + 9 * kSlice >= (1LL << 31) || offl >= (1LL << 31))
    fail(SSIM_ERR_INVALID_ARGUMENT, "mesh too large for 32-bit blocked-ELL offsets");
  auto valpos = [&](int v, int ku) -> int {  // value base of upper slot ku of row v
    const int s = v / kSlice, lane = v % kSlice;
    return 9 * c->hoffu[s] + ku * 9 * kSlice + lane;
  };
  // padding: upper slots → (v, zero values); lower references → (v, a zero block after the last slice)
  std::vector<int> hucol(std::max<long long>(offu, 1), 0);
  std::vector<int2> hlref(std::max<long long>(offl, 1), make_int2(0, 0));
  const long long zero_block = 9 * offu;
  for (int s = 0; s < c->nslices; ++s)
    for (int l = 0; l < kSlice; ++l) {
      const int v = std::min(s * kSlice + l, nv - 1);
      for (int k = 0; k < c->hwu[s]; ++k) hucol[c->hoffu[s] + k * kSlice + l] = v;
      for (int k = 0; k < hwl[s]; ++k) hlref[hoffl[s] + k * kSlice + l] = make_int2(v, (int)(zero_block + l));
    }
  for (int v = 0; v < nv; ++v) {
    const int s = v / kSlice, lane = v % kSlice;
    const int kd = c->kdiag[v];
    for (int k = kd; k < len[v]; ++k)
      hucol[c->hoffu[s] + c->bslot[c->brow[v] + k] * kSlice + lane] = c->bcol[c->brow[v] + k];
    for (int k = 0; k < kd; ++k) {
      const int w = c->bcol[c->brow[v] + k];
      hlref[hoffl[s] + c->bslot[c->brow[v] + k] * kSlice + lane] =
          make_int2(w, valpos(w, c->bslot[c->brow[w] + slot(w, v)]));
    }
  }

  // upper blocks (r ≤ c) in row-major order, with the value base of (r,c)
  std::vector<int> ubptr(nv + 1, 0);
  for (int v = 0; v < nv; ++v) ubptr[v + 1] = ubptr[v] + (len[v] - c->kdiag[v]);
  c->nub = ubptr[nv];
  std::vector<int2> hub_rc(std::max(c->nub, 1));
  std::vector<int> hub_pos(std::max(c->nub, 1));
  for (int v = 0; v < nv; ++v)
    for (int k = c->kdiag[v]; k < len[v]; ++k) {
      const int ub = ubptr[v] + (k - c->kdiag[v]);
      hub_rc[ub] = make_int2(v, c->bcol[c->brow[v] + k]);
      hub_pos[ub] = valpos(v, c->bslot[c->brow[v] + k]);
    }

  // Hessian contributions: 10 local pairs (v ≤ w) per tet, sorted by (upper block, tet, pair)
  std::vector<int> hcount(c->nub + 1, 0);
  std::vector<int> hub_of(10 * (size_t)nt, -1);
  for (int e = 0; e < nt; ++e)
    if (cells[4 * e] >= 0)
      for (int w = 0; w < 4; ++w)
        for (int v = 0; v <= w; ++v) {
          const int a = cells[4 * e + v], b = cells[4 * e + w];
          const int r = std::min(a, b), cc = std::max(a, b);
          const int ub = ubptr[r] + (slot(r, cc) - c->kdiag[r]);
          hub_of[10 * (size_t)e + w * (w + 1) / 2 + v] = ub;
          hcount[ub + 1]++;
        }
  for (int u = 0; u < c->nub; ++u) hcount[u + 1] += hcount[u];
  std::vector<int> hperm(10 * (size_t)nt, -1);
  {
    std::vector<int> pos(hcount.begin(), hcount.end() - 1);
    for (size_t k = 0; k < 10 * (size_t)nt; ++k)
      if (hub_of[k] >= 0) hperm[k] = pos[hub_of[k]]++;
  }
  // gradient contributions: 4 per tet, sorted by (vertex, tet, corner)
  std::vector<int> gseg(nv + 1, 0);
  for (int e = 0; e < nt; ++e)
    if (cells[4 * e] >= 0)
      for (int i = 0; i < 4; ++i) gseg[cells[4 * e + i] + 1]++;
  for (int v = 0; v < nv; ++v) gseg[v + 1] += gseg[v];
  std::vector<int> gperm(4 * (size_t)nt, -1);
  {
    std::vector<int> pos(gseg.begin(), gseg.end() - 1);
    for (int e = 0; e < nt; ++e)
      if (cells[4 * e] >= 0)
        for (int i = 0; i < 4; ++i) gperm[4 * e + i] = pos[cells[4 * e + i]]++;
  }
  hipStream_t s = c->stream;
  c->wu.upload(c->hwu.data(), ns, s);
  c->offu.upload(c->hoffu.data(), ns, s);
  c->ucol.upload(hucol, s);
  c->wl.upload(hwl.data(), ns, s);
  c->offl.upload(hoffl.data(), ns, s);
  c->lref.upload(hlref, s);
  c->ub_rc.upload(hub_rc, s);
  c->ub_pos.upload(hub_pos, s);
  c->hseg.upload(hcount, s);
  c->hperm.upload(hperm, s);
  c->gseg.upload(gseg, s);
  c->gperm.upload(gperm, s);
  c->val.alloc(9 * (size_t)offu + 9 * kSlice);  // + the zero block of the padded lower references
  HIPCHK(hipMemsetAsync(c->val.p, 0, c->val.count * sizeof(double), s));
  c->gcontrib.alloc(12 * (size_t)std::max(nt, 1));
  c->hcontrib.alloc(90 * (size_t)std::max(nt, 1));
  sync(c);  // host staging vectors go out of scope
}

void upload_material(ssim_ctx* c) {
  std::vector<double> lam(c->nt), mu(c->nt);
  if (c->per_tet) {
    lam = c->hlam;
    mu = c->hmu;
  } else {
    // compute_lambda / compute_mu: elast/basic_elast.hpp:11-20
    const double l = c->youngs * c->poisson / ((1 + c->poisson) * (1 - 2 * c->poisson));
    const double m = c->youngs / (2 * (1 + c->poisson));
    std::fill(lam.begin(), lam.end(), l);
    std::fill(mu.begin(), mu.end(), m);
  }
  c->lam.upload(lam, c->stream);
  c->mu.upload(mu, c->stream);
  sync(c);
}

}  // namespace

// reset(): finite_elements/time_step.hpp:368-494
namespace {
void set_precond(ssim_ctx* c, const double* values);
void build_pd(ssim_ctx* c);
}  // namespace

void ssim_hip::do_reset(ssim_ctx* c) {
  check_material(c->youngs, c->poisson, c->density);
  if (0.5 - c->poisson < 1e-4)
    fprintf(stderr, "Warning: got extremely high Poisson's ratio, may cause numerical instability. %lf\n", c->poisson);
  hipStream_t s = c->stream;
  const int n = c->n;
  c->rho_reset = c->density;
  // vertex buffers: inertia, next, forces zeroed; the others kept if already allocated (:384-413)
  for (int w = 0; w < SSIM_ARRAY_COUNT; ++w) {
    const bool keep = (w == SSIM_ARRAY_EXT_ACCEL || w == SSIM_ARRAY_DBC_VALUES || w == SSIM_ARRAY_PREV_DEFORM ||
                       w == SSIM_ARRAY_DEFORMATION || w == SSIM_ARRAY_VELOCITY) &&
                      c->arr[w].p && c->arr[w].count == (size_t)n;
    if (!keep) {
      c->arr[w].alloc(n);
      HIPCHK(hipMemsetAsync(c->arr[w].p, 0, sizeof(double) * std::max(n, 1), s));
    }
  }
  if (!(c->dof.p && c->dof.count == (size_t)n)) {
    c->dof.alloc(n);
    HIPCHK(hipMemsetAsync(c->dof.p, 0, std::max(n, 1), s));
  }
  for (DBuf<double>* b : {&c->utry, &c->gtry, &c->dir, &c->r, &c->z, &c->p, &c->q, &c->tmp, &c->dinv}) {
    b->alloc(n);
    HIPCHK(hipMemsetAsync(b->p, 0, sizeof(double) * std::max(n, 1), s));
  }
  c->hrest.release();
  c->B.alloc(9 * (size_t)std::max(c->nt, 1));
  c->V.alloc(std::max(c->nt, 1));
  upload_material(c);
  launch_rest_shape(c->X.p, c->cells.p, c->nt, c->B.p, c->V.p, s);
  HIPCHK(hipGetLastError());
  build_topology(c);
  // the SpMV grid has one workgroup per 64-row slice (or per 16 rows with 4 lanes per row), rounded up to a
  // multiple of 8 (XCD remapping)
  c->pmax = std::max<size_t>((size_t)max_partials(std::max({c->nt, c->n, c->nub, 1})), 4 * (size_t)c->nslices) + 8;
  c->partials.alloc(8 * c->pmax);
  // the projective-dynamics matrix is built at reset (projective_dynamics.hpp:51-87), for that stepper
  c->pd_val.release();
  if (c->stepper == SSIM_STEPPER_LBFGS_PD) build_pd(c);
  // the preconditioner factor is rebuilt on the new pattern and set to the identity (fem_cuda.cu:175-201)
  for (ssim_hip::DBuf<int>* b : {&c->pre_brow, &c->pre_bcol, &c->pre_tpos}) b->release();
  if (c->pre_val.p || c->stepper == SSIM_STEPPER_NCG_EXT_PREC) set_precond(c, nullptr);
  c->drop_graphs();  // buffers may have moved
  c->pcg.alloc(1);
  c->pcg_it.alloc(1);
  c->ev.alloc(1);
  HIPCHK(hipMemsetAsync(c->pcg.p, 0, sizeof(PcgScalars), s));
  HIPCHK(hipMemsetAsync(c->pcg_it.p, 0, sizeof(PcgIter), s));
  HIPCHK(hipMemsetAsync(c->ev.p, 0, sizeof(EvalScalars), s));
  // total mass Σ ρV (= asum(M.values)/3, time_step.hpp:490)
  std::vector<double> hV(c->nt);
  if (c->nt) HIPCHK(hipMemcpyAsync(hV.data(), c->V.p, sizeof(double) * c->nt, hipMemcpyDeviceToHost, s));
  sync(c);
  double mass = 0;
  for (int e = 0; e < c->nt; ++e) mass += c->density * hV[e];
  c->total_mass = mass;
  c->tol_abs = c->rtol * mass;
  c->has_reset = true;
}

namespace {

struct EvalOut {
  double E, gg;
};

// update_energy_and_gradients at deformation u → gradient into g; returns E and ‖g‖²
EvalOut eval_energy_grad(ssim_ctx* c, const double* u, double* g) {
  hipStream_t s = c->stream;
  const double dt2 = c->dt * c->dt;
  int ne = 0, ng = 0;
  timed(c, SSIM_KERNEL_ELEM_GRAD, [&] {
    ne = launch_elem_grad(c->mesh(), u, c->a(SSIM_ARRAY_INERTIA), dt2, c->gperm.p, c->gcontrib.p, c->part_e(),
                          c->prm.fp32_elements != 0, s);
  });
  timed(c, SSIM_KERNEL_GRAD_REDUCE, [&] {
    ng = launch_grad_reduce(c->nv, c->gseg.p, c->gcontrib.p, c->dof.p, g, c->part_g(), s);
  });
  launch_finalize(c->part_e(), ne, 1, c->part_g(), ng, &c->ev.p->energy, s);  // energy, gg
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->pin, c->ev.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  sync(c);
  c->stats.energy_evals++;
  return EvalOut{c->pin[0], c->pin[1]};
}

double device_dot(ssim_ctx* c, const double* a, const double* b) {
  launch_dot(c->n, a, b, c->part_d(), &c->ev.p->dot, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->pin, &c->ev.p->dot, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  sync(c);
  return c->pin[0];
}

}  // namespace

// update_hessian at deformation u into the blocked-ELL values and the Jacobi diagonal
void ssim_hip::assemble_hessian(ssim_ctx* c, const double* u, int make_spsd, int filter_dbc, double dt2,
                                double* values) {
  hipStream_t s = c->stream;
  timed(c, SSIM_KERNEL_ELEM_HESS, [&] {
    launch_elem_hess(c->mesh(), u, dt2, make_spsd, c->hperm.p, c->hcontrib.p, c->prm.fp32_elements != 0, s);
  });
  timed(c, SSIM_KERNEL_HESS_REDUCE, [&] {
    launch_hess_reduce(c->nub, c->hseg.p, c->hcontrib.p, c->ub_rc.p, c->ub_pos.p, c->dof.p, filter_dbc, values,
                       c->dinv.p, nullptr, s);
  });
  HIPCHK(hipGetLastError());
  c->stats.hessian_evals++;
}

namespace {

// Enqueues `count` PCG iterations (kernels A and B) on stream s.
void enqueue_pcg_iterations(ssim_ctx* c, const BellView& A, const double* dinv, double* x, int nb, int na, int count,
                            hipStream_t s) {
  for (int j = 0; j < count; ++j) {
    launch_pcg_spmv(A, c->z.p, c->p.p, c->q.p, c->part_pb(0), c->part_pb(1), nb, c->pcg_it.p, c->pcg.p,
                    c->part_pa(), x, s);
    launch_pcg_update(c->n, dinv, c->r.p, c->z.p, c->q.p, c->part_pa(), na, c->part_pb(0),
                      c->part_pb(1), c->pcg_it.p, c->pcg.p, s);
  }
}

// The graph of `count` iterations on matrix A writing x, captured on first use (buffers are fixed until the
// next reset, which drops the cache).
hipGraphExec_t pcg_chunk_graph(ssim_ctx* c, const BellView& A, const double* dinv, double* x, int nb, int na,
                               int count) {
  for (const auto& g : c->pcg_graphs)
    if (g.count == count && g.x == x && g.val == A.val) return g.exec;
  if (!c->cap_stream) HIPCHK(hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
  HIPCHK(hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal));
  hipGraph_t graph = nullptr;
  try {
    enqueue_pcg_iterations(c, A, dinv, x, nb, na, count, c->cap_stream);
  } catch (...) {
    (void)hipStreamEndCapture(c->cap_stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  HIPCHK(hipStreamEndCapture(c->cap_stream, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) fail(SSIM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  c->pcg_graphs.push_back({count, x, A.val, exec});
  return exec;
}

// Jacobi-PCG: x = A⁻¹ (sign·b) for the blocked-ELL values `val` with inverse diagonal `dinv` (default: the
// current Hessian).  Returns iterations.
int pcg_solve(ssim_ctx* c, const double* b, double sign, double* x, double rtol, int maxit,
              const double* val = nullptr, const double* dinv = nullptr) {
  hipStream_t s = c->stream;
  if (!val) val = c->val.p;
  if (!dinv) dinv = c->dinv.p;
  const BellView A = c->bell(val);
  const int nb = launch_pcg_init(c->n, b, sign, dinv, x, c->r.p, c->z.p, rtol, maxit, c->pcg.p, c->pcg_it.p,
                                 c->part_pb(0), s);
  HIPCHK(hipGetLastError());
  for (auto& e : c->poll_ev)
    if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  const int na = pcg_spmv_grid(A);
  // Chunks of iterations are enqueued one ahead of the done-flag poll, so the GPU never drains while the host
  // reads a flag; kernels enqueued after convergence return at entry.  The first iteration of a chunk is launched
  // on its own (it carries the sampled kernel timers); the others are replayed from a captured graph.
  int* flags = reinterpret_cast<int*>(c->pin);  // two slots of 4 ints: done, iters, maxit, breakdown
  auto enqueue = [&](int count, int slot) {
    timed_ext(c, SSIM_KERNEL_PCG_SPMV, [&](hipEvent_t e0, hipEvent_t e1) {
      launch_pcg_spmv(A, c->z.p, c->p.p, c->q.p, c->part_pb(0), c->part_pb(1), nb, c->pcg_it.p, c->pcg.p,
                      c->part_pa(), x, s, e0, e1);
    });
    timed_ext(c, SSIM_KERNEL_PCG_UPDATE, [&](hipEvent_t e0, hipEvent_t e1) {
      launch_pcg_update(c->n, dinv, c->r.p, c->z.p, c->q.p, c->part_pa(), na, c->part_pb(0),
                        c->part_pb(1), c->pcg_it.p, c->pcg.p, s, e0, e1);
    });
    if (count > 1) {
      if (c->pcg_graph)
        HIPCHK(hipGraphLaunch(pcg_chunk_graph(c, A, dinv, x, nb, na, count - 1), s));
      else
        enqueue_pcg_iterations(c, A, dinv, x, nb, na, count - 1, s);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(flags + 4 * slot, &c->pcg.p->done, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(c->poll_ev[slot], s));
  };
  int chunk = 8, launched = 0, slot = 0;
  enqueue(chunk, slot);
  launched += chunk;
  for (;;) {
    const int next = std::min(chunk * 2, std::max(8, c->pcg_chunk_max));
    enqueue(next, slot ^ 1);
    launched += next;
    HIPCHK(hipEventSynchronize(c->poll_ev[slot]));
    if (c->timing) harvest_timers(c);
    const int* f = flags + 4 * slot;
    if (f[0]) {
      if (f[3] == 1) fail(SSIM_ERR_NUMERIC, "PCG breakdown: pᵀAp ≤ 0 (matrix not positive definite)");
      if (f[3] == 2) fail(SSIM_ERR_NUMERIC, "PCG breakdown: NaN residual");
      return f[1];
    }
    if (launched > maxit + 2 * std::max(8, c->pcg_chunk_max) + 16) fail(SSIM_ERR_STATE, "internal: PCG did not stop");
    chunk = next;
    slot ^= 1;
  }
}

// Newton with backtracking Armijo line search: stepper/backward_euler.hpp:15-33 (tol = threshold·mass,
// Armijo 0.3) with the iteration cap of the benchmark configuration.
void newton_solve(ssim_ctx* c) {
  const int n = c->n;
  hipStream_t s = c->stream;
  double* u = c->a(SSIM_ARRAY_NEXT_DEFORM);
  double* g = c->a(SSIM_ARRAY_FORCES);
  EvalOut cur = eval_energy_grad(c, u, g);
  int it = 0;
  for (; it < c->prm.max_newton; ++it) {
    if (!std::isfinite(cur.E) || !std::isfinite(cur.gg)) fail(SSIM_ERR_NUMERIC, "non-finite energy or gradient");
    if (std::sqrt(cur.gg) < c->tol_abs) {
      c->stats.converged = 1;
      break;
    }
    assemble_hessian(c, u, 1, 1, c->dt * c->dt, c->val.p);
    c->stats.pcg_iterations += pcg_solve(c, g, -1.0, c->dir.p, c->prm.pcg_rtol, c->prm.pcg_maxit);
    const double gd = device_dot(c, g, c->dir.p);
    double alpha = 1.0;
    EvalOut trial{0, 0};
    bool ok = false;
    for (int ls = 0; ls < c->prm.ls_max; ++ls) {
      launch_axpy_to(n, u, alpha, c->dir.p, c->utry.p, s);
      trial = eval_energy_grad(c, c->utry.p, c->gtry.p);
      if (trial.E <= cur.E + c->prm.ls_c * alpha * gd) {
        ok = true;
        break;
      }
      alpha *= c->prm.ls_beta;
    }
    if (!ok) c->stats.ls_failures++;
    HIPCHK(hipMemcpyAsync(u, c->utry.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(g, c->gtry.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    cur = trial;
  }
  c->stats.newton_iterations = it;
  // polish: full Newton steps with the unprojected Hessian (positive definite at the converged minimum, so the
  // steps converge quadratically), no line search
  for (int k = 0; k < c->prm.newton_polish; ++k) {
    assemble_hessian(c, u, 0, 1, c->dt * c->dt, c->val.p);
    c->stats.pcg_iterations += pcg_solve(c, g, -1.0, c->dir.p, c->prm.pcg_rtol, c->prm.pcg_maxit);
    launch_axpy_to(n, u, 1.0, c->dir.p, c->utry.p, s);
    HIPCHK(hipMemcpyAsync(u, c->utry.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    eval_energy_grad(c, u, g);
  }
}

// Projective-dynamics matrix of time_step_solver_pd::reset_impl (stepper/projective_dynamics.hpp:51-87):
// M + dt²·κ·K on every dof, K the P1 Laplacian, Dirichlet rows and columns set to the identity, with κ the
// approximated diffusivity of the reference's uniform_elasticity(), which builds the material from (E, ν) in the
// (λ, μ) slots (finite_elements/time_step.hpp:356): for stable Neo-Hookean λ + 5μ/6 + 2·(4μ/3)
// (elast/stable_neohookean.hpp:109-114), κ = E + 3.5ν; for linear elasticity λ + 2μ (elast/basic_elast.hpp:225),
// κ = E + 2ν.
double pd_diffusivity(ssim_material_t model, double lam, double mu) {
  if (model == SSIM_MATERIAL_LINEAR) return lam + 2.0 * mu;
  return (lam + 5.0 / 6.0 * mu) + 2.0 * (4.0 / 3.0 * mu);
}

void build_pd(ssim_ctx* c) {
  hipStream_t s = c->stream;
  c->pd_val.alloc(c->val.count);
  c->pd_dinv.alloc(std::max(c->n, 1));
  c->pd_q.alloc(std::max(c->n, 1));
  HIPCHK(hipMemsetAsync(c->pd_val.p, 0, sizeof(double) * c->pd_val.count, s));  // padding slots stay zero
  const double kappa = pd_diffusivity(c->model, c->youngs, c->poisson);
  launch_elem_pd(c->mesh(), c->dt * c->dt * kappa, c->hperm.p, c->hcontrib.p, s);
  launch_hess_reduce(c->nub, c->hseg.p, c->hcontrib.p, c->ub_rc.p, c->ub_pos.p, c->dof.p, 1, c->pd_val.p,
                     c->pd_dinv.p, nullptr, s);
  HIPCHK(hipGetLastError());
  sync(c);
}

// H0 of the L-BFGS stepper: diagonal of the SPD-projected, Dirichlet-filtered Hessian at rest (u = 0)
void ensure_hrest(ssim_ctx* c) {
  if (c->hrest.p) return;
  const int n = c->n;
  hipStream_t s = c->stream;
  c->hrest.alloc(n);
  HIPCHK(hipMemsetAsync(c->tmp.p, 0, sizeof(double) * n, s));
  const double dt2 = c->dt * c->dt;
  launch_elem_hess(c->mesh(), c->tmp.p, dt2, 1, c->hperm.p, c->hcontrib.p, c->prm.fp32_elements != 0, s);
  launch_hess_reduce(c->nub, c->hseg.p, c->hcontrib.p, c->ub_rc.p, c->ub_pos.p, c->dof.p, 1, c->val.p, c->dinv.p,
                     c->hrest.p, s);
  HIPCHK(hipGetLastError());
}

// The same L-BFGS iteration with the two-loop recursion carried out on Gram matrices of the stored pairs (the
// vector-free formulation of Chen, Wang, Zhou & Luo, NIPS 2014, with a diagonal H0 = D): with SY_ij = s_iᵀy_j,
// YDY_ij = y_iᵀDy_j, SG_i = s_iᵀg, YDG_i = y_iᵀDg,
//   a_i = ρ_i (SG_i − Σ_{j newer} a_j SY_ij)                                  (newest to oldest)
//   b_i = ρ_i (YDG_i − Σ_j a_j YDY_ij + Σ_{j older} c_j SY_ji),  c_j = a_j − b_j   (oldest to newest)
//   r = D(g − Σ_j a_j y_j) + Σ_j c_j s_j,    gᵀr = gᵀDg − Σ_j a_j YDG_j + Σ_j c_j SG_j,
// which is the recursion of lbfgs_solve term by term.  Per iteration one pass over the stored vectors computes
// the new Gram entries (s_jᵀg, y_jᵀDg, gᵀDg and the newest pair's row and column), the host runs the scalar
// recursion, and one pass forms r.  The 2·kLbfgsSlots history vectors are zeroed once, so unused slots hold finite
// values.  If a line search fails, the rest of the solve computes the directions vector by vector on the same
// history (the recursion of lbfgs_solve).
void lbfgs_solve_gram(ssim_ctx* c) {
  const int n = c->n, m = std::max(1, c->prm.lbfgs_m);
  constexpr int M = kLbfgsSlots;
  hipStream_t s = c->stream;
  if (c->lb_S.count < (size_t)M * n || c->lb_Y.count < (size_t)M * n) {
    c->lb_S.alloc((size_t)M * n);
    c->lb_Y.alloc((size_t)M * n);
    HIPCHK(hipMemsetAsync(c->lb_S.p, 0, sizeof(double) * M * (size_t)n, s));
    HIPCHK(hipMemsetAsync(c->lb_Y.p, 0, sizeof(double) * M * (size_t)n, s));
  }
  c->lb_part.alloc((size_t)kLbfgsDots * lbfgs_gram_blocks(n));
  c->lb_scal.alloc(std::max<size_t>(c->lb_scal.count, kLbfgsDots + 2 * M + 8));
  double* av = c->lb_scal.p + kLbfgsDots;  // device two-loop coefficients of the vector-by-vector recursion
  double* bv = av + M;
  bool vector_mode = false;
  std::vector<double> SY((size_t)M * M, 0.0), YDY((size_t)M * M, 0.0), rho(M, 0.0), gram(kLbfgsDots);
  double* u = c->a(SSIM_ARRAY_NEXT_DEFORM);
  double* g = c->a(SSIM_ARRAY_FORCES);
  double* d = c->dir.p;  // r ≈ H⁻¹g; the search direction is −r
  EvalOut cur = eval_energy_grad(c, u, g);
  int it = 0, hist = 0, head = 0, fresh = -1;  // fresh: slot of the pair written by the last line search
  auto accept = [&](double ys) {
    if (!(ys > 1e-300)) return false;  // as lbfgs_solve: a pair with yᵀs ≤ 0 is dropped
    rho[fresh] = 1.0 / ys;
    head = (head + 1) % m;
    if (hist < m) hist++;
    return true;
  };
  for (; it < c->prm.lbfgs_maxit; ++it) {
    if (!std::isfinite(cur.E) || !std::isfinite(cur.gg)) fail(SSIM_ERR_NUMERIC, "non-finite energy or gradient");
    if (std::sqrt(cur.gg) < c->tol_abs) {
      c->stats.converged = 1;
      break;
    }
    double gd = 0.0;
    if (vector_mode) {
      if (fresh >= 0)
        accept(device_dot(c, c->lb_Y.p + (size_t)fresh * n, c->lb_S.p + (size_t)fresh * n));
      fresh = -1;
      HIPCHK(hipMemcpyAsync(c->q.p, g, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      for (int k = 0; k < hist; ++k) {
        const int idx = (head - 1 - k + m) % m;
        launch_dot(n, c->lb_S.p + (size_t)idx * n, c->q.p, c->part_d(), av + idx, s);
        launch_axpy_coef(n, -rho[idx], av + idx, 0.0, nullptr, c->lb_Y.p + (size_t)idx * n, c->q.p, s);
      }
      launch_scale_div(n, c->q.p, c->hrest.p, d, s);
      for (int k = hist - 1; k >= 0; --k) {
        const int idx = (head - 1 - k + m) % m;
        launch_dot(n, c->lb_Y.p + (size_t)idx * n, d, c->part_d(), bv + idx, s);
        launch_axpy_coef(n, rho[idx], av + idx, rho[idx], bv + idx, c->lb_S.p + (size_t)idx * n, d, s);
      }
      HIPCHK(hipGetLastError());
      gd = -device_dot(c, g, d);
    } else {
      launch_lbfgs_gram(n, c->lb_S.p, c->lb_Y.p, g, c->hrest.p, fresh, c->lb_part.p, c->lb_scal.p, s);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(gram.data(), c->lb_scal.p, sizeof(double) * kLbfgsDots, hipMemcpyDeviceToHost, s));
      sync(c);
      const double* SG = gram.data();
      const double* YDG = SG + M;
      const double GDG = SG[2 * M];
      if (fresh >= 0) {
        const double* SYk = SG + 2 * M + 1;  // s_jᵀy_k
        const double* SkY = SG + 3 * M + 1;  // s_kᵀy_j
        const double* YDYk = SG + 4 * M + 1;
        for (int j = 0; j < M; ++j) {  // row and column of the slot; used only if the pair is kept
          SY[(size_t)j * M + fresh] = SYk[j];
          SY[(size_t)fresh * M + j] = SkY[j];
          YDY[(size_t)fresh * M + j] = YDY[(size_t)j * M + fresh] = YDYk[j];
        }
        accept(SkY[fresh]);
        fresh = -1;
      }
      // scalar two-loop recursion over the history, slot of the k-th newest pair: (head − 1 − k) mod m
      LbfgsCoef cf{};
      double a[M] = {}, b[M] = {};
      for (int k = 0; k < hist; ++k) {
        const int i = (head - 1 - k + m) % m;
        double t = SG[i];
        for (int q = 0; q < k; ++q) {
          const int j = (head - 1 - q + m) % m;
          t -= a[j] * SY[(size_t)i * M + j];
        }
        a[i] = rho[i] * t;
      }
      for (int k = hist - 1; k >= 0; --k) {
        const int i = (head - 1 - k + m) % m;
        double t = YDG[i];
        for (int q = 0; q < hist; ++q) {
          const int j = (head - 1 - q + m) % m;
          t -= a[j] * YDY[(size_t)i * M + j];
        }
        for (int q = k + 1; q < hist; ++q) {  // older pairs, already updated
          const int j = (head - 1 - q + m) % m;
          t += (a[j] - b[j]) * SY[(size_t)j * M + i];
        }
        b[i] = rho[i] * t;
      }
      double gr = GDG;
      for (int k = 0; k < hist; ++k) {
        const int i = (head - 1 - k + m) % m;
        cf.a[i] = a[i];
        cf.c[i] = a[i] - b[i];
        cf.mask |= 1u << i;
        gr += -a[i] * YDG[i] + (a[i] - b[i]) * SG[i];
      }
      launch_lbfgs_dir(n, c->lb_S.p, c->lb_Y.p, g, c->hrest.p, cf, d, s);
      HIPCHK(hipGetLastError());
      gd = -gr;
    }
    double alpha = 1.0;
    EvalOut trial{0, 0};
    bool ok = false;
    for (int ls = 0; ls < c->prm.lbfgs_ls_max; ++ls) {
      launch_axpy_to(n, u, -alpha, d, c->utry.p, s);
      trial = eval_energy_grad(c, c->utry.p, c->gtry.p);
      if (trial.E <= cur.E + c->prm.lbfgs_c * alpha * gd) {
        ok = true;
        break;
      }
      alpha *= c->prm.ls_beta;
    }
    if (!ok) {
      c->stats.ls_failures++;
      vector_mode = true;
    }
    launch_lbfgs_shift(n, c->utry.p, c->gtry.p, u, g, c->lb_S.p + (size_t)head * n, c->lb_Y.p + (size_t)head * n, s);
    HIPCHK(hipGetLastError());
    fresh = head;
    cur = trial;
  }
  c->stats.newton_iterations = it;
}

// L-BFGS two-loop recursion (Nocedal & Wright, Alg. 7.4) with H0 = diag(H_rest)⁻¹ and backtracking Armijo;
// stepper/backward_euler.hpp:48-60 (tol = threshold·mass).  The two-loop coefficients stay on the device;
// ρ_i = 1/(y_iᵀs_i) is read back once per accepted pair.  With `pd` the middle step solves with the
// projective-dynamics matrix instead (time_step_solver_pd, stepper/projective_dynamics.hpp:31-50: a direct solve
// in the reference, Jacobi-PCG to a relative residual of kPdRtol here).
// With `h0` it solves with the SPD-projected, Dirichlet-filtered Hessian at the step's starting point instead
// (time_step_solver_h0, stepper/projective_dynamics.hpp:91-148: update_hessian(false, true) and a factorisation per
// step; the reference binds no Python class for it).
constexpr double kPdRtol = 1e-12;
void lbfgs_solve(ssim_ctx* c, bool pd, bool h0 = false) {
  const int n = c->n, m = std::max(1, c->prm.lbfgs_m);
  hipStream_t s = c->stream;
  if (pd && !c->pd_val.p) build_pd(c);
  if (h0) {
    c->pd_q.alloc(std::max(n, 1));
    assemble_hessian(c, c->a(SSIM_ARRAY_NEXT_DEFORM), 1, 1, c->dt * c->dt, c->val.p);
  }
  const bool solve = pd || h0;
  double* qv = solve ? c->pd_q.p : c->q.p;  // the solve's PCG uses c->q itself
  if (!solve) ensure_hrest(c);
  if (!solve && c->lbfgs_gram && m <= kLbfgsSlots) {
    lbfgs_solve_gram(c);
    return;
  }
  c->lb_S.alloc((size_t)m * n);
  c->lb_Y.alloc((size_t)m * n);
  c->lb_scal.alloc(2 * (size_t)m + 8);  // a[m], b[m]
  double* av = c->lb_scal.p;
  double* bv = av + m;
  std::vector<double> rho_h(m, 0.0);
  double* u = c->a(SSIM_ARRAY_NEXT_DEFORM);
  double* g = c->a(SSIM_ARRAY_FORCES);
  double* d = c->dir.p;  // holds r ≈ H⁻¹g; the search direction is −r
  EvalOut cur = eval_energy_grad(c, u, g);
  int it = 0, hist = 0, head = 0;
  for (; it < c->prm.lbfgs_maxit; ++it) {
    if (!std::isfinite(cur.E) || !std::isfinite(cur.gg)) fail(SSIM_ERR_NUMERIC, "non-finite energy or gradient");
    if (std::sqrt(cur.gg) < c->tol_abs) {
      c->stats.converged = 1;
      break;
    }
    HIPCHK(hipMemcpyAsync(qv, g, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    for (int k = 0; k < hist; ++k) {
      const int idx = (head - 1 - k + m) % m;
      // a_i = ρ_i s_iᵀq ; q −= a_i y_i
      launch_dot(n, c->lb_S.p + (size_t)idx * n, qv, c->part_d(), av + idx, s);
      launch_axpy_coef(n, -rho_h[idx], av + idx, 0.0, nullptr, c->lb_Y.p + (size_t)idx * n, qv, s);
    }
    if (pd)  // r = A_pd⁻¹ q
      c->stats.pcg_iterations += pcg_solve(c, qv, 1.0, d, kPdRtol, c->prm.pcg_maxit, c->pd_val.p, c->pd_dinv.p);
    else if (h0)  // r = H(u_start)⁻¹ q
      c->stats.pcg_iterations += pcg_solve(c, qv, 1.0, d, kPdRtol, c->prm.pcg_maxit);
    else
      launch_scale_div(n, qv, c->hrest.p, d, s);  // r = H0 q
    for (int k = hist - 1; k >= 0; --k) {
      const int idx = (head - 1 - k + m) % m;
      // b = ρ_i y_iᵀr ; r += s_i (a_i − b)
      launch_dot(n, c->lb_Y.p + (size_t)idx * n, d, c->part_d(), bv + idx, s);
      launch_axpy_coef(n, rho_h[idx], av + idx, rho_h[idx], bv + idx, c->lb_S.p + (size_t)idx * n, d, s);
    }
    HIPCHK(hipGetLastError());
    const double gd = -device_dot(c, g, d);
    double alpha = 1.0;
    EvalOut trial{0, 0};
    bool ok = false;
    for (int ls = 0; ls < c->prm.lbfgs_ls_max; ++ls) {
      launch_axpy_to(n, u, -alpha, d, c->utry.p, s);
      trial = eval_energy_grad(c, c->utry.p, c->gtry.p);
      if (trial.E <= cur.E + c->prm.lbfgs_c * alpha * gd) {
        ok = true;
        break;
      }
      alpha *= c->prm.ls_beta;
    }
    if (!ok) c->stats.ls_failures++;
    double* sv = c->lb_S.p + (size_t)head * n;
    double* yv = c->lb_Y.p + (size_t)head * n;
    launch_sub(n, c->utry.p, u, sv, s);
    launch_sub(n, c->gtry.p, g, yv, s);
    const double ys = device_dot(c, yv, sv);
    if (ys > 1e-300) {
      rho_h[head] = 1.0 / ys;
      head = (head + 1) % m;
      if (hist < m) hist++;
    }
    HIPCHK(hipMemcpyAsync(u, c->utry.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(g, c->gtry.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    cur = trial;
  }
  c->stats.newton_iterations = it;
}

// The preconditioner factor L of the externally preconditioned NCG lives on the block adjacency (the system
// matrix's pattern, fem_cuda.cu:175-201) as block CSR, with tpos[k] the index of the mirrored block (w, v) of
// block k = (v, w) for the transposed product.  `values` null: the identity.
void set_precond(ssim_ctx* c, const double* values) {
  const int nv = c->nv;
  const long long nb = c->brow[nv];
  if (!c->pre_brow.p) {
    std::vector<int> tpos(std::max<long long>(nb, 1));
    for (int v = 0; v < nv; ++v)
      for (int k = c->brow[v]; k < c->brow[v + 1]; ++k) {
        const int w = c->bcol[k];
        const int* b = c->bcol.data() + c->brow[w];
        const int* e = c->bcol.data() + c->brow[w + 1];
        tpos[k] = c->brow[w] + (int)(std::lower_bound(b, e, v) - b);
      }
    c->pre_brow.upload(c->brow.data(), nv + 1, c->stream);
    c->pre_bcol.upload(c->bcol.data(), std::max<long long>(nb, 1), c->stream);
    c->pre_tpos.upload(tpos, c->stream);
    for (ssim_hip::DBuf<double>* b : {&c->pre_a, &c->pre_b, &c->pre_pg}) b->alloc(std::max(c->n, 1));
    sync(c);
  }
  // user CSR order: scalar row 3v+b holds, for each neighbour k of v in column order, the entries (b, 0..2);
  // it starts at 9·brow[v] + 3·b·len(v)
  std::vector<double> hv(9 * (size_t)std::max<long long>(nb, 1), 0.0);
  for (int v = 0; v < nv; ++v) {
    const int len = c->brow[v + 1] - c->brow[v];
    for (int k = 0; k < len; ++k) {
      double* B = hv.data() + 9 * (size_t)(c->brow[v] + k);
      if (!values) {
        if (c->bcol[c->brow[v] + k] == v) B[0] = B[4] = B[8] = 1.0;
        continue;
      }
      for (int b = 0; b < 3; ++b)
        for (int d = 0; d < 3; ++d) B[3 * b + d] = values[9 * (size_t)c->brow[v] + 3 * (size_t)b * len + 3 * k + d];
    }
  }
  c->pre_val.upload(hv, c->stream);
  sync(c);
}

// z = L·(Lᵀ·g): a = Lᵀ g, out = L a
void apply_precond(ssim_ctx* c, const double* g, double* a, double* out) {
  launch_bsr_spmv(c->nv, c->pre_brow.p, c->pre_bcol.p, c->pre_tpos.p, c->pre_val.p, g, a, true, c->stream);
  launch_bsr_spmv(c->nv, c->pre_brow.p, c->pre_bcol.p, c->pre_tpos.p, c->pre_val.p, a, out, false, c->stream);
  HIPCHK(hipGetLastError());
}

// Nonlinear CG (stepper/nonlinear_cg.hpp:9-37: Dai-Kou strategy, learning rate 1e-2, backtracking line search,
// tol = threshold·mass), the published DK+ method: β = max(gₖ₊₁ᵀy/dᵀy − ‖y‖²·gₖ₊₁ᵀd/(dᵀy)², ½·gₖ₊₁ᵀd/‖d‖²),
// d = −gₖ₊₁ + βd, restarts with d = −g when dᵀy ≤ 0 or d is not a descent direction; backtracking Armijo with the
// L-BFGS constants from a first trial of 1e-2, then αₖ₋₁·gₖ₋₁ᵀdₖ₋₁/gₖᵀdₖ capped at 1.
// With the external preconditioner P = L·Lᵀ (stepper/nonlinear_cg.hpp:41-97) the preconditioned gradient P·g
// replaces g in the direction and the first term of β reads gₖ₊₁ᵀPy/dᵀy − yᵀPy·gₖ₊₁ᵀd/(dᵀy)² (the Euclidean
// floor is kept); with L = I every quantity is the unpreconditioned one.
void ncg_solve(ssim_ctx* c, bool precond) {
  const int n = c->n;
  hipStream_t s = c->stream;
  double* u = c->a(SSIM_ARRAY_NEXT_DEFORM);
  double* g = c->a(SSIM_ARRAY_FORCES);
  double* d = c->dir.p;
  c->lb_scal.alloc(std::max<size_t>(c->lb_scal.count, 16));
  double* dots = c->lb_scal.p;  // five reductions of one iteration
  // their partials (5 per vector block) fit the dot-product slot: pmax ≥ 4·slices = n/48 > 5·(n/1024 + 1)
  if (5 * (size_t)pcg_vector_blocks(n) > c->pmax) fail(SSIM_ERR_STATE, "internal: partial buffer too small");
  if (precond && !c->pre_val.p) set_precond(c, nullptr);
  double* pa = precond ? c->pre_a.p : nullptr;  // Lᵀ·g of the current gradient
  double* pb = precond ? c->pre_b.p : nullptr;  // Lᵀ·g of the trial gradient
  double* pg = precond ? c->pre_pg.p : nullptr;  // P·g
  EvalOut cur = eval_energy_grad(c, u, g);
  if (precond) {
    apply_precond(c, g, pa, pg);
    launch_axpby(n, -1.0, pg, 0.0, d, s);
  } else {
    launch_axpby(n, -1.0, g, 0.0, d, s);
  }
  double alpha_prev = 1e-2, gd_prev = 0.0;
  int it = 0;
  for (; it < c->prm.lbfgs_maxit; ++it) {
    if (!std::isfinite(cur.E) || !std::isfinite(cur.gg)) fail(SSIM_ERR_NUMERIC, "non-finite energy or gradient");
    if (std::sqrt(cur.gg) < c->tol_abs) {
      c->stats.converged = 1;
      break;
    }
    double gd = device_dot(c, g, d);
    if (!(gd < 0)) {  // not a descent direction: restart along −P·g
      launch_axpby(n, -1.0, precond ? pg : g, 0.0, d, s);
      gd = precond ? -device_dot(c, pa, pa) : -cur.gg;
    }
    double alpha = it == 0 ? 1e-2 : alpha_prev * gd_prev / gd;
    if (!(alpha > 0) || alpha > 1.0) alpha = 1.0;
    EvalOut trial{0, 0};
    bool ok = false;
    for (int ls = 0; ls < c->prm.lbfgs_ls_max; ++ls) {
      launch_axpy_to(n, u, alpha, d, c->utry.p, s);
      trial = eval_energy_grad(c, c->utry.p, c->gtry.p);
      if (trial.E <= cur.E + c->prm.lbfgs_c * alpha * gd) {
        ok = true;
        break;
      }
      alpha *= c->prm.ls_beta;
    }
    if (!ok) c->stats.ls_failures++;
    if (precond) apply_precond(c, c->gtry.p, pb, pg);
    launch_ncg_dots(n, c->gtry.p, g, d, precond ? pb : c->gtry.p, precond ? pa : g, c->part_d(), dots, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->pin, dots, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
    sync(c);
    const double dy = c->pin[0], gy = c->pin[1], yy = c->pin[2], gnd = c->pin[3], dd = c->pin[4];
    double beta = 0.0;
    if (dy > 0) beta = std::max(gy / dy - yy * gnd / (dy * dy), 0.5 * gnd / dd);
    launch_axpby(n, -1.0, precond ? pg : c->gtry.p, beta, d, s);
    HIPCHK(hipMemcpyAsync(u, c->utry.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(g, c->gtry.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    if (precond) std::swap(pa, pb);
    cur = trial;
    alpha_prev = alpha;
    gd_prev = gd;
  }
  c->stats.newton_iterations = it;
}

// Forward Euler (time_step_solver_forward_euler, time_step.hpp:717-736): one evaluation of the forces at the
// predictor held in next_deform, then next_deform += ½·dt²·forces.  The forces array is the DBC-filtered
// gradient of the incremental potential (time_step.hpp:577-624) and the update adds it with a positive sign,
// exactly as the reference does; Dirichlet dofs keep their targets because their gradient entries are zero.
void forward_euler_solve(ssim_ctx* c) {
  double* u = c->a(SSIM_ARRAY_NEXT_DEFORM);
  double* f = c->a(SSIM_ARRAY_FORCES);
  const EvalOut cur = eval_energy_grad(c, u, f);
  if (!std::isfinite(cur.E) || !std::isfinite(cur.gg)) fail(SSIM_ERR_NUMERIC, "non-finite energy or gradient");
  launch_axpby(c->n, 0.5 * c->dt * c->dt, f, 1.0, u, c->stream);
  HIPCHK(hipGetLastError());
  c->stats.newton_iterations = 1;
  c->stats.converged = 1;
}

ssim_status_t guard(ssim_ctx* c) {
  if (!c) {
    g_err = "null context";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  return SSIM_OK;
}

void require_reset(ssim_ctx* c) {
  if (!c->has_reset) fail(SSIM_ERR_STATE, "context not reset");
}

}  // namespace

void ssim_hip::edit_dirichlet(ssim_ctx* c, const int* verts, int count, const double* targets, bool mark) {
  if (count < 0 || (count > 0 && !verts)) fail(SSIM_ERR_INVALID_ARGUMENT, "bad vertex list");
  for (int i = 0; i < count; ++i)
    if (verts[i] < 0 || verts[i] >= c->nv) fail(SSIM_ERR_INVALID_ARGUMENT, "vertex index out of range");
  std::vector<unsigned char> dof(c->n);
  std::vector<double> dbc(c->n);
  hipStream_t s = c->stream;
  HIPCHK(hipMemcpyAsync(dof.data(), c->dof.p, c->n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(dbc.data(), c->a(SSIM_ARRAY_DBC_VALUES), sizeof(double) * c->n, hipMemcpyDeviceToHost, s));
  sync(c);
  for (int i = 0; i < count; ++i)
    for (int j = 0; j < 3; ++j) {
      dof[3 * (size_t)verts[i] + j] = mark ? 1 : 0;
      dbc[3 * (size_t)verts[i] + j] = mark ? (targets ? targets[3 * (size_t)i + j] : 0.0) : 0.0;
    }
  HIPCHK(hipMemcpyAsync(c->dof.p, dof.data(), c->n, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->a(SSIM_ARRAY_DBC_VALUES), dbc.data(), sizeof(double) * c->n, hipMemcpyHostToDevice, s));
  sync(c);
  c->hrest.release();
}

// ======================================================================================================
// extern "C" ABI

extern "C" {

const char* ssim_last_error(void) { return g_err.c_str(); }

ssim_status_t ssim_create(ssim_ctx** out, const double* vertices, int nv, const int* cells, int nt, double time_step,
                          double youngs, double poisson, double density) {
  if (!out) {
    g_err = "out is null";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  ssim_ctx* c = nullptr;
  const ssim_status_t st = run([&] {
    if (nv < 0 || nt < 0) fail(SSIM_ERR_INVALID_ARGUMENT, "negative mesh size");
    if ((nv > 0 && !vertices) || (nt > 0 && !cells)) fail(SSIM_ERR_INVALID_ARGUMENT, "null mesh arrays");
    if (!(time_step > 0)) fail(SSIM_ERR_INVALID_ARGUMENT, "Time step must be positive.");
    if ((long long)3 * nv >= (1LL << 31) || (long long)12 * nt >= (1LL << 31))
      fail(SSIM_ERR_INVALID_ARGUMENT, "mesh too large for 32-bit indices");
    for (long long k = 0; k < 4LL * nt; ++k)
      if (cells[k] < 0 || cells[k] >= nv)
        fail(SSIM_ERR_INVALID_ARGUMENT, "cell " + std::to_string(k / 4) + " references vertex " + std::to_string(cells[k]) +
                                            " outside [0, " + std::to_string(nv) + ")");
    check_material(youngs, poisson, density);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) fail(SSIM_ERR_HIP, "no HIP device available");
    c = new ssim_ctx();
    c->nv = nv;
    c->nt = nt;
    c->n = 3 * nv;
    c->dt = time_step;
    c->youngs = youngs;
    c->poisson = poisson;
    c->density = density;
    default_params(c->prm);
    check_material(youngs, poisson, density);
    c->hX.assign(vertices, vertices + 3 * (size_t)nv);
    c->hcells.assign(cells, cells + 4 * (size_t)nt);
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->pin), 64 * sizeof(double), hipHostMallocDefault));
    c->X.upload(c->hX, c->stream);
    c->cells.upload(reinterpret_cast<const int4*>(c->hcells.data()), nt, c->stream);
    do_reset(c);  // the constructor resets (time_step.hpp:260-265)
  });
  if (st != SSIM_OK) {
    delete c;
    return st;
  }
  *out = c;
  return SSIM_OK;
}

ssim_status_t ssim_destroy(ssim_ctx* ctx) {
  if (!ctx) return SSIM_OK;
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
  return SSIM_OK;
}

ssim_status_t ssim_set_stream(ssim_ctx* ctx, void* hip_stream) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    sync(ctx);
    ctx->stream = reinterpret_cast<hipStream_t>(hip_stream);
  });
}

ssim_status_t ssim_set_time_step(ssim_ctx* ctx, double dt) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (!(dt > 0)) fail(SSIM_ERR_INVALID_ARGUMENT, "Time step must be positive.");
    ctx->dt = dt;
  });
}
ssim_status_t ssim_set_youngs(ssim_ctx* ctx, double youngs) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (!(youngs > 0)) fail(SSIM_ERR_INVALID_ARGUMENT, "Young's modulus must be positive.");
    ctx->youngs = youngs;
  });
}
ssim_status_t ssim_set_poisson(ssim_ctx* ctx, double poisson) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (!(poisson > 0 && poisson < 0.5)) fail(SSIM_ERR_INVALID_ARGUMENT, "Poisson's ratio must be in (0, 0.5).");
    ctx->poisson = poisson;
  });
}
ssim_status_t ssim_set_density(ssim_ctx* ctx, double density) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (!(density > 0)) fail(SSIM_ERR_INVALID_ARGUMENT, "Density must be positive.");
    ctx->density = density;
  });
}
ssim_status_t ssim_set_damping(ssim_ctx* ctx, double damping) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (!(0 <= damping && damping < 1)) fail(SSIM_ERR_INVALID_ARGUMENT, "Damping must be in [0, 1).");
    ctx->damping = damping;
  });
}
ssim_status_t ssim_set_threshold(ssim_ctx* ctx, double rtol) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    // stored only; the absolute tolerance is rtol·mass recomputed at reset (time_step.hpp:343-347, :482-492)
    if (!(0 < rtol && rtol < 1)) fail(SSIM_ERR_INVALID_ARGUMENT, "Threshold must be positive.");
    ctx->rtol = rtol;
  });
}
ssim_status_t ssim_get_physical_params(const ssim_ctx* ctx, double* time_step, double* youngs, double* poisson,
                                       double* density, double* damping, double* threshold) {
  if (!ctx) {
    g_err = "null context";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  if (time_step) *time_step = ctx->dt;
  if (youngs) *youngs = ctx->youngs;
  if (poisson) *poisson = ctx->poisson;
  if (density) *density = ctx->density;
  if (damping) *damping = ctx->damping;
  if (threshold) *threshold = ctx->rtol;
  return SSIM_OK;
}
ssim_status_t ssim_set_material_per_tet(ssim_ctx* ctx, const double* lambda, const double* mu) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (!lambda || !mu) {
      ctx->per_tet = false;
      ctx->hlam.clear();
      ctx->hmu.clear();
    } else {
      for (int e = 0; e < ctx->nt; ++e)
        if (!(lambda[e] > 0) || !(mu[e] > 0) || !std::isfinite(lambda[e]) || !std::isfinite(mu[e]))
          fail(SSIM_ERR_INVALID_ARGUMENT, "per-tet Lamé parameters must be positive and finite (tet " + std::to_string(e) + ")");
      ctx->per_tet = true;
      ctx->hlam.assign(lambda, lambda + ctx->nt);
      ctx->hmu.assign(mu, mu + ctx->nt);
    }
    upload_material(ctx);
    ctx->hrest.release();
  });
}
ssim_status_t ssim_material_eval(int dim, ssim_material_t model, long long n, const double* F, const double* lambda,
                                 const double* mu, int make_spsd, double* psi, double* P, double* H, void* stream) {
  return run([&] {
    if (dim != 2 && dim != 3) fail(SSIM_ERR_INVALID_ARGUMENT, "dim must be 2 or 3");
    if (model != SSIM_MATERIAL_STABLE_NEOHOOKEAN && model != SSIM_MATERIAL_LINEAR)
      fail(SSIM_ERR_INVALID_ARGUMENT, "unknown material model");
    if (n < 0) fail(SSIM_ERR_INVALID_ARGUMENT, "n must be non-negative");
    if (n > 0 && (!F || !lambda || !mu)) fail(SSIM_ERR_INVALID_ARGUMENT, "null input array");
    if (make_spsd && dim != 3) fail(SSIM_ERR_INVALID_ARGUMENT, "the SPD projection is defined for dim 3 only");
    if (n == 0) return;
    launch_material_eval(dim, model, n, F, lambda, mu, make_spsd, psi, P, H, static_cast<hipStream_t>(stream));
    HIPCHK(hipGetLastError());
  });
}

ssim_status_t ssim_set_material_model(ssim_ctx* ctx, ssim_material_t model) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  if (model != SSIM_MATERIAL_STABLE_NEOHOOKEAN && model != SSIM_MATERIAL_LINEAR) {
    g_err = "unknown material model";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  return run([&] {
    if (ctx->model != model) ctx->hrest.release();
    ctx->model = model;
  });
}
ssim_status_t ssim_get_material_model(const ssim_ctx* ctx, ssim_material_t* model) {
  if (!ctx || !model) {
    g_err = "null argument";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  *model = ctx->model;
  return SSIM_OK;
}
ssim_status_t ssim_set_solver_params(ssim_ctx* ctx, const ssim_solver_params* p) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (!p) fail(SSIM_ERR_INVALID_ARGUMENT, "params is null");
    if (p->max_newton < 0 || p->pcg_maxit < 0 || !(p->pcg_rtol >= 0) || !(p->ls_c > 0 && p->ls_c < 1) ||
        !(p->ls_beta > 0 && p->ls_beta < 1) || p->ls_max < 1 || p->lbfgs_m < 1 || p->lbfgs_maxit < 0 ||
        !(p->lbfgs_c > 0 && p->lbfgs_c < 1) || p->lbfgs_ls_max < 1 || p->newton_polish < 0 || p->newton_polish > 16)
      fail(SSIM_ERR_INVALID_ARGUMENT, "solver parameter out of range");
    const bool fp32_changed = (p->fp32_elements != 0) != (ctx->prm.fp32_elements != 0);
    ctx->prm = *p;
    if (fp32_changed) ctx->hrest.release();
  });
}
ssim_status_t ssim_get_solver_params(const ssim_ctx* ctx, ssim_solver_params* p) {
  if (!ctx || !p) {
    g_err = "null argument";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  *p = ctx->prm;
  return SSIM_OK;
}
ssim_status_t ssim_set_stepper(ssim_ctx* ctx, ssim_stepper_t stepper) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  if (stepper != SSIM_STEPPER_NEWTON && stepper != SSIM_STEPPER_LBFGS && stepper != SSIM_STEPPER_NCG &&
      stepper != SSIM_STEPPER_FORWARD_EULER && stepper != SSIM_STEPPER_NCG_EXT_PREC &&
      stepper != SSIM_STEPPER_LBFGS_PD && stepper != SSIM_STEPPER_LBFGS_H0) {
    g_err = "unknown stepper";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  ctx->stepper = stepper;
  return SSIM_OK;
}

ssim_status_t ssim_mark_dirichlet(ssim_ctx* ctx, const int* verts, int count, const double* targets) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] { edit_dirichlet(ctx, verts, count, targets, true); });
}
ssim_status_t ssim_mark_general(ssim_ctx* ctx, const int* verts, int count) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] { edit_dirichlet(ctx, verts, count, nullptr, false); });
}

ssim_status_t ssim_add_ext_force_dof(ssim_ctx* ctx, int dof_idx, double value) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (!(-3 < dof_idx && dof_idx < 3)) fail(SSIM_ERR_INVALID_ARGUMENT, "|dof_idx| must be less than PhysicalDim.");
    if (dof_idx < 0) dof_idx += 3;
    std::vector<double> acc(ctx->n);
    HIPCHK(hipMemcpyAsync(acc.data(), ctx->a(SSIM_ARRAY_EXT_ACCEL), sizeof(double) * ctx->n, hipMemcpyDeviceToHost,
                          ctx->stream));
    sync(ctx);
    for (int v = 0; v < ctx->nv; ++v) acc[3 * (size_t)v + dof_idx] += value;
    HIPCHK(hipMemcpyAsync(ctx->a(SSIM_ARRAY_EXT_ACCEL), acc.data(), sizeof(double) * ctx->n, hipMemcpyHostToDevice,
                          ctx->stream));
    sync(ctx);
  });
}

ssim_status_t ssim_reset(ssim_ctx* ctx) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] { do_reset(ctx); });
}

ssim_status_t ssim_prepare_step(ssim_ctx* ctx) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    require_reset(ctx);
    if (!(0 <= ctx->damping && ctx->damping < 1)) fail(SSIM_ERR_INVALID_ARGUMENT, "velocity damping must be in [0, 1).");
    launch_prepare(ctx->n, ctx->a(SSIM_ARRAY_DEFORMATION), ctx->a(SSIM_ARRAY_VELOCITY), ctx->a(SSIM_ARRAY_EXT_ACCEL),
                   ctx->dof.p, ctx->a(SSIM_ARRAY_DBC_VALUES), ctx->dt, ctx->a(SSIM_ARRAY_INERTIA),
                   ctx->a(SSIM_ARRAY_NEXT_DEFORM), ctx->stream);
    HIPCHK(hipGetLastError());
  });
}

ssim_status_t ssim_compute_step(ssim_ctx* ctx, double* residual_out) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    require_reset(ctx);
    ctx->stats = ssim_step_stats{};
    if (ctx->stepper == SSIM_STEPPER_NEWTON)
      newton_solve(ctx);
    else if (ctx->stepper == SSIM_STEPPER_LBFGS || ctx->stepper == SSIM_STEPPER_LBFGS_PD ||
             ctx->stepper == SSIM_STEPPER_LBFGS_H0)
      lbfgs_solve(ctx, ctx->stepper == SSIM_STEPPER_LBFGS_PD, ctx->stepper == SSIM_STEPPER_LBFGS_H0);
    else if (ctx->stepper == SSIM_STEPPER_NCG || ctx->stepper == SSIM_STEPPER_NCG_EXT_PREC)
      ncg_solve(ctx, ctx->stepper == SSIM_STEPPER_NCG_EXT_PREC);
    else
      forward_euler_solve(ctx);
    // compute_step: final update_energy_and_gradients and ‖forces‖ (time_step.hpp:520-528)
    const EvalOut fin = eval_energy_grad(ctx, ctx->a(SSIM_ARRAY_NEXT_DEFORM), ctx->a(SSIM_ARRAY_FORCES));
    ctx->stats.grad_norm = std::sqrt(fin.gg);
    ctx->stats.energy = fin.E;
    if (!std::isfinite(fin.E) || !std::isfinite(fin.gg)) fail(SSIM_ERR_NUMERIC, "non-finite energy or gradient");
    if (residual_out) *residual_out = ctx->stats.grad_norm;
  });
}

ssim_status_t ssim_step_next(ssim_ctx* ctx) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    require_reset(ctx);
    launch_step_next(ctx->n, ctx->a(SSIM_ARRAY_NEXT_DEFORM), ctx->a(SSIM_ARRAY_DEFORMATION),
                     ctx->a(SSIM_ARRAY_PREV_DEFORM), ctx->a(SSIM_ARRAY_VELOCITY), ctx->dt, ctx->damping, ctx->stream);
    HIPCHK(hipGetLastError());
  });
}

ssim_status_t ssim_step(ssim_ctx* ctx, double* residual_out) {
  ssim_status_t st = ssim_prepare_step(ctx);
  if (st) return st;
  st = ssim_compute_step(ctx, residual_out);
  if (st) return st;
  return ssim_step_next(ctx);
}

ssim_status_t ssim_restore(ssim_ctx* ctx) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    require_reset(ctx);
    launch_restore(ctx->n, ctx->dof.p, ctx->a(SSIM_ARRAY_DBC_VALUES), ctx->a(SSIM_ARRAY_DEFORMATION),
                   ctx->a(SSIM_ARRAY_PREV_DEFORM), ctx->a(SSIM_ARRAY_VELOCITY), ctx->dt, ctx->stream);
    HIPCHK(hipGetLastError());
  });
}

ssim_status_t ssim_get_step_stats(const ssim_ctx* ctx, ssim_step_stats* stats) {
  if (!ctx || !stats) {
    g_err = "null argument";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  *stats = ctx->stats;
  return SSIM_OK;
}

ssim_status_t ssim_update_energy_and_gradients(ssim_ctx* ctx, double* energy_out) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    require_reset(ctx);
    const EvalOut r = eval_energy_grad(ctx, ctx->a(SSIM_ARRAY_NEXT_DEFORM), ctx->a(SSIM_ARRAY_FORCES));
    if (energy_out) *energy_out = r.E;
  });
}

ssim_status_t ssim_update_hessian(ssim_ctx* ctx, int make_spsd, int filter_dbc) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    require_reset(ctx);
    assemble_hessian(ctx, ctx->a(SSIM_ARRAY_NEXT_DEFORM), make_spsd ? 1 : 0, filter_dbc ? 1 : 0, ctx->dt * ctx->dt,
                     ctx->val.p);
    sync(ctx);
  });
}

ssim_status_t ssim_check_convergence(ssim_ctx* ctx, int* converged_out) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    require_reset(ctx);
    launch_dot(ctx->n, ctx->a(SSIM_ARRAY_FORCES), ctx->a(SSIM_ARRAY_FORCES), ctx->part_d(), &ctx->ev.p->dot,
               ctx->stream);
    HIPCHK(hipMemcpyAsync(ctx->pin, &ctx->ev.p->dot, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    if (converged_out) *converged_out = std::sqrt(ctx->pin[0]) < ctx->tol_abs;
  });
}

ssim_status_t ssim_get_array(ssim_ctx* ctx, ssim_array_t which, double* host_out) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (which < 0 || which >= SSIM_ARRAY_COUNT || !host_out) fail(SSIM_ERR_INVALID_ARGUMENT, "bad array request");
    HIPCHK(hipMemcpyAsync(host_out, ctx->a(which), sizeof(double) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
  });
}

ssim_status_t ssim_set_array(ssim_ctx* ctx, ssim_array_t which, const double* host_in) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (which < 0 || which >= SSIM_ARRAY_COUNT || !host_in) fail(SSIM_ERR_INVALID_ARGUMENT, "bad array request");
    HIPCHK(hipMemcpyAsync(ctx->a(which), host_in, sizeof(double) * ctx->n, hipMemcpyHostToDevice, ctx->stream));
    sync(ctx);
  });
}

ssim_status_t ssim_get_array_device(ssim_ctx* ctx, ssim_array_t which, double* dst) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (which < 0 || which >= SSIM_ARRAY_COUNT || !dst) fail(SSIM_ERR_INVALID_ARGUMENT, "bad array request");
    HIPCHK(hipMemcpyAsync(dst, ctx->a(which), sizeof(double) * ctx->n, hipMemcpyDeviceToDevice, ctx->stream));
    sync(ctx);
  });
}

ssim_status_t ssim_set_array_device(ssim_ctx* ctx, ssim_array_t which, const double* src) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (which < 0 || which >= SSIM_ARRAY_COUNT || !src) fail(SSIM_ERR_INVALID_ARGUMENT, "bad array request");
    HIPCHK(hipMemcpyAsync(ctx->a(which), src, sizeof(double) * ctx->n, hipMemcpyDeviceToDevice, ctx->stream));
    sync(ctx);
  });
}

// Adjoint of one implicit step at the converged u* = next_deform.  The step solves g(u*; û, λ, μ) = 0 on the
// free dofs, g = M(u − û) + dt²·Σ_e V_e f_e(u; λ_e, μ_e) (time_step.hpp:577-624), so for a loss L(u*):
//   H l = ∂L/∂u*  (H = ∂g/∂u, the unprojected Hessian with the Dirichlet identity filter; PCG),
//   ∂L/∂û = M l (zero on Dirichlet dofs),   ∂L/∂θ = −(∂g/∂θ)ᵀ l for the per-tet Lamé parameters θ.
ssim_status_t ssim_adjoint_step(ssim_ctx* ctx, const double* grad_u, double* grad_uhat, double* grad_lambda,
                                double* grad_mu, double* adjoint_out, int* pcg_iters_out) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    require_reset(ctx);
    if (!grad_u || !grad_uhat) fail(SSIM_ERR_INVALID_ARGUMENT, "grad_u and grad_uhat are required");
    ssim_ctx* c = ctx;
    hipStream_t s = c->stream;
    const double dt2 = c->dt * c->dt;
    const double* u = c->a(SSIM_ARRAY_NEXT_DEFORM);
    assemble_hessian(c, u, 0, 1, dt2, c->val.p);
    launch_mask_dbc(c->n, c->dof.p, grad_u, c->tmp.p, s);
    HIPCHK(hipGetLastError());
    const int it = pcg_solve(c, c->tmp.p, 1.0, c->dir.p, c->prm.pcg_rtol, c->prm.pcg_maxit);
    double* gl = grad_lambda;
    double* gm = grad_mu;
    if (!gl) {
      c->adj_lam.alloc(std::max(c->nt, 1));
      gl = c->adj_lam.p;
    }
    if (!gm) {
      c->adj_mu.alloc(std::max(c->nt, 1));
      gm = c->adj_mu.p;
    }
    launch_elem_adjoint(c->mesh(), u, c->dir.p, dt2, c->gperm.p, c->gcontrib.p, gl, gm, s);
    launch_grad_reduce(c->nv, c->gseg.p, c->gcontrib.p, c->dof.p, grad_uhat, c->part_g(), s);
    HIPCHK(hipGetLastError());
    if (adjoint_out)
      HIPCHK(hipMemcpyAsync(adjoint_out, c->dir.p, sizeof(double) * c->n, hipMemcpyDeviceToDevice, s));
    sync(c);
    if (pcg_iters_out) *pcg_iters_out = it;
  });
}

ssim_status_t ssim_device_array(ssim_ctx* ctx, ssim_array_t which, double** device_ptr) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  if (which < 0 || which >= SSIM_ARRAY_COUNT || !device_ptr) {
    g_err = "bad array request";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  *device_ptr = ctx->a(which);
  return SSIM_OK;
}

ssim_status_t ssim_get_dof_type(const ssim_ctx* ctx, unsigned char* host_out) {
  if (!ctx || !host_out) {
    g_err = "null argument";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  return run([&] {
    HIPCHK(hipMemcpyAsync(host_out, ctx->dof.p, ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

ssim_status_t ssim_num_vertices(const ssim_ctx* ctx, int* nv) {
  if (!ctx || !nv) return SSIM_ERR_INVALID_ARGUMENT;
  *nv = ctx->nv;
  return SSIM_OK;
}
ssim_status_t ssim_num_cells(const ssim_ctx* ctx, int* nt) {
  if (!ctx || !nt) return SSIM_ERR_INVALID_ARGUMENT;
  *nt = ctx->nt;
  return SSIM_OK;
}
ssim_status_t ssim_total_mass(const ssim_ctx* ctx, double* mass) {
  if (!ctx || !mass) return SSIM_ERR_INVALID_ARGUMENT;
  *mass = ctx->total_mass;
  return SSIM_OK;
}
ssim_status_t ssim_grad_threshold_abs(const ssim_ctx* ctx, double* tol) {
  if (!ctx || !tol) return SSIM_ERR_INVALID_ARGUMENT;
  *tol = ctx->tol_abs;
  return SSIM_OK;
}

ssim_status_t ssim_csr_nnz(const ssim_ctx* ctx, long long* nnz) {
  if (!ctx || !nnz) return SSIM_ERR_INVALID_ARGUMENT;
  *nnz = 9LL * (ctx->brow.empty() ? 0 : ctx->brow[ctx->nv]);
  return SSIM_OK;
}

ssim_status_t ssim_set_precond_matrix(ssim_ctx* ctx, const double* values, long long nnz) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    require_reset(ctx);
    if (!values) fail(SSIM_ERR_INVALID_ARGUMENT, "values is null");
    const long long want = 9LL * ctx->brow[ctx->nv];
    if (nnz != want)
      fail(SSIM_ERR_INVALID_ARGUMENT, "the preconditioner needs " + std::to_string(want) + " values (the CSR pattern of "
                                       "the system matrix), got " + std::to_string(nnz));
    set_precond(ctx, values);
  });
}

// blocked-ELL values (device) → scalar CSR in row/column order
static void export_csr(ssim_ctx* c, const double* dval, int* rowptr, int* colidx, double* values) {
  const int nv = c->nv;
  if (rowptr) {
    rowptr[0] = 0;
    for (int v = 0; v < nv; ++v) {
      const int len = 3 * (c->brow[v + 1] - c->brow[v]);
      for (int b = 0; b < 3; ++b) rowptr[3 * v + b + 1] = rowptr[3 * v + b] + len;
    }
  }
  if (colidx) {
    long long p = 0;
    for (int v = 0; v < nv; ++v)
      for (int b = 0; b < 3; ++b)
        for (int k = c->brow[v]; k < c->brow[v + 1]; ++k)
          for (int d = 0; d < 3; ++d) colidx[p++] = 3 * c->bcol[k] + d;
  }
  if (values) {
    std::vector<double> hv(9 * (size_t)std::max<long long>(c->ncolslots, 1));
    HIPCHK(hipMemcpyAsync(hv.data(), dval, sizeof(double) * hv.size(), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    auto base = [&](int v, int ku) -> size_t {
      const int s = v / kSlice, lane = v % kSlice;
      return 9 * (size_t)c->hoffu[s] + (size_t)ku * 9 * kSlice + lane;
    };
    long long p = 0;
    for (int v = 0; v < nv; ++v) {
      const int len = c->brow[v + 1] - c->brow[v];
      const int kd = c->kdiag[v];
      for (int b = 0; b < 3; ++b)
        for (int k = 0; k < len; ++k) {
          const int w = c->bcol[c->brow[v] + k];
          if (k >= kd) {
            const size_t bs = base(v, c->bslot[c->brow[v] + k]);
            for (int d = 0; d < 3; ++d) values[p++] = hv[bs + (size_t)(3 * b + d) * kSlice];
          } else {  // lower block (v, w) = transpose of the stored (w, v)
            const int* bw = c->bcol.data() + c->brow[w];
            const int kw = (int)(std::lower_bound(bw, bw + (c->brow[w + 1] - c->brow[w]), v) - bw);
            const size_t bs = base(w, c->bslot[c->brow[w] + kw]);
            for (int d = 0; d < 3; ++d) values[p++] = hv[bs + (size_t)(3 * d + b) * kSlice];
          }
        }
    }
  }
}

ssim_status_t ssim_export_hessian_csr(ssim_ctx* ctx, int* rowptr, int* colidx, double* values) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    require_reset(ctx);
    export_csr(ctx, ctx->val.p, rowptr, colidx, values);
  });
}

ssim_status_t ssim_export_pd_csr(ssim_ctx* ctx, int* rowptr, int* colidx, double* values) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    require_reset(ctx);
    if (!ctx->pd_val.p) build_pd(ctx);
    export_csr(ctx, ctx->pd_val.p, rowptr, colidx, values);
  });
}

namespace {
// the mass matrix is the Hessian assembly with dt = 0; `filtered` applies the Dirichlet identity rows and columns
// (mass_matrix_filtered_, time_step.hpp:460-466) with the current dof flags
ssim_status_t export_mass(ssim_ctx* ctx, int filtered, int* rowptr, int* colidx, double* values) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    require_reset(ctx);
    if (values) {
      ctx->massval.alloc(ctx->val.count);
      HIPCHK(hipMemsetAsync(ctx->massval.p, 0, sizeof(double) * ctx->val.count, ctx->stream));
      launch_elem_hess(ctx->mesh(), ctx->a(SSIM_ARRAY_NEXT_DEFORM), 0.0, 0, ctx->hperm.p, ctx->hcontrib.p, false,
                       ctx->stream);
      launch_hess_reduce(ctx->nub, ctx->hseg.p, ctx->hcontrib.p, ctx->ub_rc.p, ctx->ub_pos.p, ctx->dof.p, filtered,
                         ctx->massval.p, ctx->tmp.p, nullptr, ctx->stream);
      HIPCHK(hipGetLastError());
    }
    export_csr(ctx, ctx->massval.p, rowptr, colidx, values);
  });
}
}  // namespace

ssim_status_t ssim_export_mass_csr(ssim_ctx* ctx, int* rowptr, int* colidx, double* values) {
  return export_mass(ctx, 0, rowptr, colidx, values);
}

ssim_status_t ssim_export_mass_filtered_csr(ssim_ctx* ctx, int* rowptr, int* colidx, double* values) {
  return export_mass(ctx, 1, rowptr, colidx, values);
}

ssim_status_t ssim_get_rest_shape(ssim_ctx* ctx, double* dminv, double* rest_vol) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    require_reset(ctx);
    const size_t nt = ctx->nt;
    if (dminv && nt) {
      std::vector<double> soa(9 * nt);
      HIPCHK(hipMemcpyAsync(soa.data(), ctx->B.p, sizeof(double) * 9 * nt, hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      for (size_t e = 0; e < nt; ++e)
        for (int k = 0; k < 9; ++k) dminv[9 * e + k] = soa[(size_t)k * nt + e];
    }
    if (rest_vol && nt) {
      HIPCHK(hipMemcpyAsync(rest_vol, ctx->V.p, sizeof(double) * nt, hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
    }
  });
}

ssim_status_t ssim_device_mesh(const ssim_ctx* ctx, const double** vertices, const int** cells) {
  if (!ctx || (!vertices && !cells)) {
    g_err = "null argument";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  if (vertices) *vertices = ctx->X.p;
  if (cells) *cells = reinterpret_cast<const int*>(ctx->cells.p);
  return SSIM_OK;
}

ssim_status_t ssim_eval_elements(ssim_ctx* ctx, const double* u, int make_spsd, double* psi, double* P, double* H,
                                 double* K) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    require_reset(ctx);
    if (!u) fail(SSIM_ERR_INVALID_ARGUMENT, "u is null");
    const size_t nt = ctx->nt;
    HIPCHK(hipMemcpyAsync(ctx->tmp.p, u, sizeof(double) * ctx->n, hipMemcpyHostToDevice, ctx->stream));
    DBuf<double> dpsi, dP, dH, dK;
    if (psi) dpsi.alloc(nt);
    if (P) dP.alloc(9 * nt);
    if (H) dH.alloc(81 * nt);
    if (K) dK.alloc(144 * nt);
    launch_elem_eval(ctx->mesh(), ctx->tmp.p, make_spsd ? 1 : 0, dpsi.p, dP.p, dH.p, dK.p, ctx->prm.fp32_elements != 0,
                     ctx->stream);
    HIPCHK(hipGetLastError());
    if (psi) HIPCHK(hipMemcpyAsync(psi, dpsi.p, sizeof(double) * nt, hipMemcpyDeviceToHost, ctx->stream));
    if (P) HIPCHK(hipMemcpyAsync(P, dP.p, sizeof(double) * 9 * nt, hipMemcpyDeviceToHost, ctx->stream));
    if (H) HIPCHK(hipMemcpyAsync(H, dH.p, sizeof(double) * 81 * nt, hipMemcpyDeviceToHost, ctx->stream));
    if (K) HIPCHK(hipMemcpyAsync(K, dK.p, sizeof(double) * 144 * nt, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
  });
}

ssim_status_t ssim_pcg_solve(ssim_ctx* ctx, const double* b, double* x, double rtol, int maxit, int* iters_out) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    require_reset(ctx);
    if (!b || !x) fail(SSIM_ERR_INVALID_ARGUMENT, "null vector");
    HIPCHK(hipMemcpyAsync(ctx->tmp.p, b, sizeof(double) * ctx->n, hipMemcpyHostToDevice, ctx->stream));
    const int it = pcg_solve(ctx, ctx->tmp.p, 1.0, ctx->dir.p, rtol, maxit);
    HIPCHK(hipMemcpyAsync(x, ctx->dir.p, sizeof(double) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    if (iters_out) *iters_out = it;
  });
}

ssim_status_t ssim_spmv(ssim_ctx* ctx, const double* x, double* y) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    require_reset(ctx);
    if (!x || !y) fail(SSIM_ERR_INVALID_ARGUMENT, "null vector");
    HIPCHK(hipMemcpyAsync(ctx->tmp.p, x, sizeof(double) * ctx->n, hipMemcpyHostToDevice, ctx->stream));
    launch_spmv(ctx->bell(ctx->val.p), ctx->tmp.p, ctx->q.p, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(y, ctx->q.p, sizeof(double) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
  });
}

ssim_status_t ssim_set_kernel_timing(ssim_ctx* ctx, int enabled) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (auto& t : ctx->kt) {
      t.sum_ms = 0;
      t.cnt = 0;
      t.pending = false;
    }
    ctx->timing = enabled != 0;
  });
}

ssim_status_t ssim_get_kernel_timing(const ssim_ctx* ctx, ssim_kernel_t kernel, double* avg_ms, long long* samples) {
  if (!ctx || kernel < 0 || kernel >= SSIM_KERNEL_COUNT) {
    g_err = "bad kernel timing request";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  const auto& t = ctx->kt[kernel];
  if (avg_ms) *avg_ms = t.cnt ? t.sum_ms / t.cnt : 0.0;
  if (samples) *samples = t.cnt;
  return SSIM_OK;
}

ssim_status_t ssim_get_layout_info(const ssim_ctx* ctx, ssim_layout_info* info) {
  if (!ctx || !info) {
    g_err = "null argument";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  info->nv = ctx->nv;
  info->nt = ctx->nt;
  info->nonzero_blocks = ctx->brow.empty() ? 0 : ctx->brow[ctx->nv];
  info->padded_slots = ctx->ncolslots;
  info->lower_slots = ctx->nlowslots;
  info->upper_blocks = ctx->nub;
  info->slices = ctx->nslices;
  return SSIM_OK;
}

ssim_status_t ssim_set_tuning(ssim_ctx* ctx, ssim_tune_t key, int value) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    sync(ctx);
    switch (key) {
      case SSIM_TUNE_XCD_REMAP:
        ctx->xcd_remap = value != 0;
        ctx->drop_graphs();
        break;
      case SSIM_TUNE_PCG_CHUNK_MAX:
        if (value < 1 || value > 4096) fail(SSIM_ERR_INVALID_ARGUMENT, "PCG chunk must be in [1, 4096]");
        ctx->pcg_chunk_max = value;
        break;
      case SSIM_TUNE_SPMV_LANES:
        if (value != 0 && value != 1 && value != 4) fail(SSIM_ERR_INVALID_ARGUMENT, "SpMV variant must be 0, 1 or 4");
        if (value == 0 && !ctx->fixed_layout)
          fail(SSIM_ERR_INVALID_ARGUMENT, "the fixed-width SpMV needs at most 8 upper and 7 lower blocks per row");
        ctx->spmv_lanes = value;
        ctx->spmv_lanes_default = false;
        ctx->drop_graphs();
        break;
      case SSIM_TUNE_PCG_GRAPH:
        ctx->pcg_graph = value != 0;
        break;
      case SSIM_TUNE_LBFGS_GRAM:
        ctx->lbfgs_gram = value != 0;
        break;
      default:
        fail(SSIM_ERR_INVALID_ARGUMENT, "unknown tuning key");
    }
  });
}

ssim_status_t ssim_get_tuning(const ssim_ctx* ctx, ssim_tune_t key, int* value) {
  if (!ctx || !value) {
    g_err = "null argument";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  switch (key) {
    case SSIM_TUNE_XCD_REMAP:
      *value = ctx->xcd_remap;
      return SSIM_OK;
    case SSIM_TUNE_PCG_CHUNK_MAX:
      *value = ctx->pcg_chunk_max;
      return SSIM_OK;
    case SSIM_TUNE_SPMV_LANES:
      *value = ctx->spmv_lanes;
      return SSIM_OK;
    case SSIM_TUNE_PCG_GRAPH:
      *value = ctx->pcg_graph;
      return SSIM_OK;
    case SSIM_TUNE_LBFGS_GRAM:
      *value = ctx->lbfgs_gram;
      return SSIM_OK;
    default:
      g_err = "unknown tuning key";
      return SSIM_ERR_INVALID_ARGUMENT;
  }
}

}  // extern "C"
