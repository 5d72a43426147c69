// batch.hip -- host side of batched stepping: many independent meshes of uneven size advanced together.
//
// The reference steps one mesh per time-step object (finite_elements/time_step.hpp); a batch here holds many
// such meshes in one context and runs each phase of their Newton solves -- energy and gradient, Hessian
// assembly, PCG iterations, line-search trials -- as one set of launches over all meshes that still take part.
// Every mesh follows exactly the single-mesh algorithm (backward_euler.hpp:15-33 with the same threshold,
// Armijo constant, line-search and PCG limits); only the scheduling differs: Newton iterations advance in
// lock step, a converged mesh sits out (its workgroups return at entry), and each line-search trial evaluates
// only the meshes still searching.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "context.h"

using namespace ssim_hip;

struct ssim_batch {
  ssim_ctx c;  // the concatenated, padded mesh
  int nmesh = 0;
  std::vector<int> nv_m, nt_m;      // real sizes
  std::vector<int> vbase, tbase;    // first vertex / tet of each mesh in the concatenation
  std::vector<int> hvwg_off, htwg_off;
  int nvwg = 0, ntwg = 0;
  DBuf<int> vwg_mesh, twg_mesh, vwg_off, twg_off, tend;
  DBuf<int> vcoff, nvm, tcoff;  // compact (caller's) vertex / tet offsets and vertex counts per mesh
  long long nv_compact = 0, nt_compact = 0;
  DBuf<int> mask_all, mask_newton, mask_ls;
  DBuf<PcgScalars> pcg;
  DBuf<EvalScalars> ev;
  DBuf<PcgIter> iter;
  DBuf<double> alpha;
  std::vector<double> mass, tol;
  std::vector<ssim_step_stats> stats;  // per mesh, last compute_step
  ssim_step_stats total{};             // batch: lock-step iteration counts
  unsigned char* pin = nullptr;        // pinned staging for per-mesh scalars
  size_t pin_bytes = 0;
  // compacted PCG grids: slice and vertex-workgroup lists of the meshes still iterating, double-buffered so that
  // a list is rebuilt while the chunk enqueued ahead of the poll still reads the other one
  DBuf<int4> slist[2], wlist[2];  // {slice or vertex workgroup, mesh, mesh's first vertex workgroup, count}
  int4* pin_list[2] = {nullptr, nullptr};  // pinned staging: slices, then vertex workgroups
  int ns_list[2] = {0, 0}, nw_list[2] = {0, 0};
  int compact_pct = 70;  // rebuild when the active slices fall to this percentage of the current list
  // captured chunks of batched PCG iterations, keyed by everything their launches embed
  struct Graph {
    int count, buf, ns, nw;
    const double* x;
    const int* mask;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  int graph_mode = 0;  // replay chunks from graphs (measured: the same speed as plain launches)
  void drop_graphs() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
  }

  ~ssim_batch() {
    drop_graphs();
    if (pin) (void)hipHostFree(pin);
    for (int4* q : pin_list)
      if (q) (void)hipHostFree(q);
  }
  BatchView view(const DBuf<int>& mask) const {
    BatchView b;
    b.nmesh = nmesh;
    b.vwg_mesh = vwg_mesh.p;
    b.twg_mesh = twg_mesh.p;
    b.vwg_off = vwg_off.p;
    b.twg_off = twg_off.p;
    b.tend = tend.p;
    b.mask = mask.p;
    return b;
  }
  EvalScalars* pin_ev() { return reinterpret_cast<EvalScalars*>(pin); }
  PcgScalars* pin_pcg(int slot) {
    return reinterpret_cast<PcgScalars*>(pin + sizeof(EvalScalars) * (size_t)nmesh) + (size_t)slot * nmesh;
  }
};

namespace {

ssim_status_t bguard(const ssim_batch* b) {
  if (!b) {
    g_err = "null batch";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  return SSIM_OK;
}

void check_mesh(const ssim_batch* b, int mesh, bool allow_all) {
  if (mesh < (allow_all ? -1 : 0) || mesh >= b->nmesh)
    fail(SSIM_ERR_INVALID_ARGUMENT, "mesh index " + std::to_string(mesh) + " outside [0, " + std::to_string(b->nmesh) +
                                        ")");
}

void upload_mask(ssim_batch* b, DBuf<int>& dst, const std::vector<int>& h) {
  HIPCHK(hipMemcpyAsync(dst.p, h.data(), sizeof(int) * b->nmesh, hipMemcpyHostToDevice, b->c.stream));
}

// per-mesh masses and thresholds; padding vertices Dirichlet-fixed (after do_reset)
void batch_reset(ssim_batch* b) {
  ssim_ctx* c = &b->c;
  b->drop_graphs();
  do_reset(c);
  std::vector<double> hV(c->nt);
  HIPCHK(hipMemcpyAsync(hV.data(), c->V.p, sizeof(double) * c->nt, hipMemcpyDeviceToHost, c->stream));
  std::vector<unsigned char> dof(c->n);
  HIPCHK(hipMemcpyAsync(dof.data(), c->dof.p, c->n, hipMemcpyDeviceToHost, c->stream));
  sync(c);
  b->mass.assign(b->nmesh, 0.0);
  b->tol.assign(b->nmesh, 0.0);
  for (int m = 0; m < b->nmesh; ++m) {
    double s = 0;
    for (int e = b->tbase[m]; e < b->tbase[m] + b->nt_m[m]; ++e) s += c->rho_reset * hV[e];
    b->mass[m] = s;
    b->tol[m] = c->rtol * s;
    const int pad0 = b->vbase[m] + b->nv_m[m], pad1 = kBatchRows * b->hvwg_off[m + 1];
    for (int v = pad0; v < pad1; ++v)
      for (int a = 0; a < 3; ++a) dof[3 * (size_t)v + a] = 1;
  }
  HIPCHK(hipMemcpyAsync(c->dof.p, dof.data(), c->n, hipMemcpyHostToDevice, c->stream));
  b->pcg.alloc(b->nmesh);
  b->ev.alloc(b->nmesh);
  b->iter.alloc(1);
  b->alpha.alloc(b->nmesh);
  HIPCHK(hipMemsetAsync(b->pcg.p, 0, sizeof(PcgScalars) * b->nmesh, c->stream));
  HIPCHK(hipMemsetAsync(b->ev.p, 0, sizeof(EvalScalars) * b->nmesh, c->stream));
  HIPCHK(hipMemsetAsync(b->iter.p, 0, sizeof(PcgIter), c->stream));
  sync(c);
}

// energy and gradient of the masked meshes at u → g; per-mesh (E, ‖g‖²) into the pinned mirror
void beval(ssim_batch* b, const DBuf<int>& mask, const double* u, double* g) {
  ssim_ctx* c = &b->c;
  hipStream_t s = c->stream;
  const BatchView bt = b->view(mask);
  timed(c, SSIM_KERNEL_ELEM_GRAD, [&] {
    launch_belem_grad_tile(c->mesh(), bt, b->ntwg, u, c->a(SSIM_ARRAY_INERTIA), c->dt * c->dt, c->gt_slot.p,
                           c->gt_off.p, c->gt_run.p, c->gpart.p, c->part_e(), c->prm.fp32_elements != 0, s);
  });
  timed(c, SSIM_KERNEL_GRAD_REDUCE, [&] {
    launch_bgrad_gather(bt, b->nvwg, c->gv_seg.p, c->gv_pos.p, c->gpart.p, c->dof.p, g, c->part_g(), s);
  });
  launch_bfinalize_eval(bt, c->part_e(), c->part_g(), b->ev.p, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b->pin_ev(), b->ev.p, sizeof(EvalScalars) * b->nmesh, hipMemcpyDeviceToHost, s));
  sync(c);
}

void bassemble(ssim_batch* b, const DBuf<int>& mask, const double* u, int make_spsd) {
  ssim_ctx* c = &b->c;
  hipStream_t s = c->stream;
  const BatchView bt = b->view(mask);
  if (c->hess_tile) {
    timed(c, SSIM_KERNEL_ELEM_HESS, [&] {
      launch_belem_hess_tile(c->mesh(), bt, b->ntwg, u, c->dt * c->dt, make_spsd, c->ht_slot.p, c->ht_off.p,
                             c->ht_run.p, c->hpart.p, c->hparts, c->prm.fp32_elements != 0, s);
    });
    timed(c, SSIM_KERNEL_HESS_REDUCE, [&] {
      launch_bhess_gather(bt, c->nub, c->hb_seg.p, c->hb_pos.p, c->hpart.p, c->hparts, c->ub_rc.p, c->ub_pos.p,
                          c->dof.p, 1, c->val.p, c->dinv.p, s);
    });
  } else {
    ensure_hcontrib(c);
    timed(c, SSIM_KERNEL_ELEM_HESS, [&] {
      launch_belem_hess(c->mesh(), bt, b->ntwg, u, c->dt * c->dt, make_spsd, c->hperm.p, c->hcontrib.p,
                        c->prm.fp32_elements != 0, s);
    });
    timed(c, SSIM_KERNEL_HESS_REDUCE, [&] {
      launch_bhess_reduce(bt, c->nub, c->hseg.p, c->hcontrib.p, c->ub_rc.p, c->ub_pos.p, c->dof.p, 1, c->val.p,
                          c->dinv.p, s);
    });
  }
  HIPCHK(hipGetLastError());
}

// Writes the slice / vertex-workgroup lists of the meshes with active[m] into list buffer `buf` (copies on the
// stream, from pinned staging that is not rewritten before the copy has run: the alternate buffer's staging is
// only reused after a later poll).
void build_lists(ssim_batch* b, const std::vector<int>& active, int buf) {
  ssim_ctx* c = &b->c;
  int4* st = b->pin_list[buf];
  int ns = 0, nw = 0;
  for (int m = 0; m < b->nmesh; ++m)
    if (active[m]) {
      const int w0 = b->hvwg_off[m], nwm = b->hvwg_off[m + 1] - w0;
      for (int w = w0; w < w0 + nwm; ++w)
        for (int q = 0; q < kBatchSlices; ++q) st[ns++] = make_int4(kBatchSlices * w + q, m, w0, nwm);
    }
  int4* wst = st + c->nslices;
  for (int m = 0; m < b->nmesh; ++m)
    if (active[m]) {
      const int w0 = b->hvwg_off[m], nwm = b->hvwg_off[m + 1] - w0;
      for (int w = w0; w < w0 + nwm; ++w) wst[nw++] = make_int4(w, m, w0, nwm);
    }
  if (ns) HIPCHK(hipMemcpyAsync(b->slist[buf].p, st, sizeof(int4) * ns, hipMemcpyHostToDevice, c->stream));
  if (nw) HIPCHK(hipMemcpyAsync(b->wlist[buf].p, wst, sizeof(int4) * nw, hipMemcpyHostToDevice, c->stream));
  b->ns_list[buf] = ns;
  b->nw_list[buf] = nw;
}

// batched Jacobi-PCG of H x = sign·rhs for the masked meshes; per-mesh iteration counts into iters
void bpcg(ssim_batch* b, const DBuf<int>& mask, const double* rhs, double sign, double* x, std::vector<int>& iters) {
  ssim_ctx* c = &b->c;
  hipStream_t s = c->stream;
  const BatchView bt = b->view(mask);
  const BellView A = c->bell(c->val.p);
  const int maxit = c->prm.pcg_maxit;
  launch_bpcg_init(bt, b->nvwg, rhs, sign, c->dinv.p, x, c->r.p, c->z.p, c->prm.pcg_rtol, maxit, b->pcg.p, b->iter.p,
                   c->part_pb(0), s);
  HIPCHK(hipGetLastError());
  for (auto& e : c->poll_ev)
    if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  // the grids start with every slice of the masked meshes
  std::vector<int> active(b->nmesh);
  HIPCHK(hipMemcpyAsync(active.data(), mask.p, sizeof(int) * b->nmesh, hipMemcpyDeviceToHost, s));
  sync(c);
  int cur = 0;
  build_lists(b, active, cur);
  auto launch_iters = [&](int count, const int4* sl, int ns, const int4* wl, int nw, hipStream_t st) {
    for (int j = 0; j < count; ++j) {
      launch_bpcg_spmv(A, sl, ns, c->z.p, c->p.p, c->q.p, c->part_pb(0), c->part_pb(1), b->pcg.p, b->iter.p,
                       c->part_pa(), st);
      launch_bpcg_update(wl, nw, c->dinv.p, x, c->r.p, c->z.p, c->p.p, c->q.p, c->part_pa(), c->part_pb(0),
                         c->part_pb(1), b->pcg.p, b->iter.p, st);
    }
  };
  // a chunk of `count` iterations over list buffer `buf` as a captured graph (captured on first use; lists are
  // rewritten in place, so a graph stays valid while its list length is unchanged)
  auto chunk_graph = [&](int count, int buf) -> hipGraphExec_t {
    const int ns = b->ns_list[buf], nw = b->nw_list[buf];
    for (const auto& g : b->graphs)
      if (g.count == count && g.buf == buf && g.ns == ns && g.nw == nw && g.x == x && g.mask == mask.p) return g.exec;
    if (b->graphs.size() >= 24) b->drop_graphs();
    if (!c->cap_stream) HIPCHK(hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
    HIPCHK(hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal));
    hipGraph_t graph = nullptr;
    try {
      launch_iters(count, b->slist[buf].p, ns, b->wlist[buf].p, nw, c->cap_stream);
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
    b->graphs.push_back({count, buf, ns, nw, x, mask.p, exec});
    return exec;
  };
  auto enqueue = [&](int count, int slot) {
    const int4* sl = b->slist[cur].p;
    const int4* wl = b->wlist[cur].p;
    const int ns = b->ns_list[cur], nw = b->nw_list[cur];
    // the chunk's first iteration is launched on its own (it carries the sampled kernel timers)
    timed(c, SSIM_KERNEL_PCG_SPMV, [&] {
      launch_bpcg_spmv(A, sl, ns, c->z.p, c->p.p, c->q.p, c->part_pb(0), c->part_pb(1), b->pcg.p, b->iter.p,
                       c->part_pa(), s);
    });
    timed(c, SSIM_KERNEL_PCG_UPDATE, [&] {
      launch_bpcg_update(wl, nw, c->dinv.p, x, c->r.p, c->z.p, c->p.p, c->q.p, c->part_pa(), c->part_pb(0),
                         c->part_pb(1), b->pcg.p, b->iter.p, s);
    });
    if (count > 1) {
      if (b->graph_mode)
        HIPCHK(hipGraphLaunch(chunk_graph(count - 1, cur), s));
      else
        launch_iters(count - 1, sl, ns, wl, nw, s);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(b->pin_pcg(slot), b->pcg.p, sizeof(PcgScalars) * b->nmesh, hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(c->poll_ev[slot], s));
  };
  if (b->ns_list[cur] == 0) {  // nothing to solve: every masked-out mesh was marked done by the init
    for (int m = 0; m < b->nmesh; ++m) iters[m] = 0;
    return;
  }
  // Chunk sizes double from 8 up to pcg_chunk_max; near the end the chunk after the one in flight covers the
  // remainder predicted for the slowest mesh still iterating, from the decrease of its ‖r‖² between its last two
  // polls (10 % + 4 margin, at least 4), as in the single-mesh solve.
  const int cmax = std::max(8, c->pcg_chunk_max);
  int launched = 0, slot = 0;
  enqueue(8, slot);
  launched += 8;
  int next = std::min(16, cmax);
  std::vector<int> k_prev(b->nmesh, -1);
  std::vector<double> rr_prev(b->nmesh, 0.0);
  for (;;) {
    enqueue(next, slot ^ 1);
    launched += next;
    wait_event(c->poll_ev[slot]);
    if (c->timing) harvest_timers(c);
    const PcgScalars* f = b->pin_pcg(slot);
    int left_slices = 0;
    double remaining = 0.0;  // predicted iterations after the chunk in flight, over the meshes still iterating
    for (int m = 0; m < b->nmesh; ++m) {
      active[m] = f[m].done ? 0 : 1;
      if (!active[m]) continue;
      left_slices += kBatchSlices * (b->hvwg_off[m + 1] - b->hvwg_off[m]);
      double rem = 1e300;
      if (k_prev[m] >= 0 && f[m].iters > k_prev[m] && f[m].rr > 0.0 && rr_prev[m] > 0.0 && f[m].tol2 > 0.0) {
        const double slope = std::log(f[m].rr / rr_prev[m]) / (f[m].iters - k_prev[m]);
        if (slope < 0.0) rem = std::log(f[m].tol2 / f[m].rr) / slope - next;
      }
      remaining = std::max(remaining, rem);
      k_prev[m] = f[m].iters;
      rr_prev[m] = f[m].rr;
    }
    if (left_slices == 0) {
      for (int m = 0; m < b->nmesh; ++m) {
        if (f[m].breakdown == 1)
          fail(SSIM_ERR_NUMERIC, "mesh " + std::to_string(m) + ": PCG breakdown: pᵀAp ≤ 0 (matrix not positive definite)");
        if (f[m].breakdown == 2) fail(SSIM_ERR_NUMERIC, "mesh " + std::to_string(m) + ": PCG breakdown: NaN residual");
        iters[m] = f[m].iters;
      }
      return;
    }
    if (launched > maxit + 2 * std::max(8, c->pcg_chunk_max) + 16) fail(SSIM_ERR_STATE, "internal: PCG did not stop");
    // compact the grids once enough meshes have finished (the chunk already enqueued keeps the old list)
    if (100LL * left_slices <= (long long)b->compact_pct * b->ns_list[cur]) {
      cur ^= 1;
      build_lists(b, active, cur);
    }
    int want = std::min(next * 2, cmax);
    if (remaining < want) want = std::max(4, (int)std::ceil(1.1 * remaining + 4.0));
    next = std::min(want, cmax);
    slot ^= 1;
  }
}

// Newton with backtracking Armijo line search, per mesh, in lock step (backward_euler.hpp:15-33)
void bnewton(ssim_batch* b) {
  ssim_ctx* c = &b->c;
  hipStream_t s = c->stream;
  const int M = b->nmesh;
  const ssim_solver_params& P = c->prm;
  double* u = c->a(SSIM_ARRAY_NEXT_DEFORM);
  double* g = c->a(SSIM_ARRAY_FORCES);
  std::vector<double> E(M), gg(M), gd(M), alpha(M), Et(M), ggt(M);
  std::vector<int> act(M, 1), search(M, 0), iters(M, 0);
  for (auto& st : b->stats) st = ssim_step_stats{};
  b->total = ssim_step_stats{};
  beval(b, b->mask_all, u, g);
  for (int m = 0; m < M; ++m) {
    E[m] = b->pin_ev()[m].energy;
    gg[m] = b->pin_ev()[m].gg;
    b->stats[m].energy_evals++;
  }
  b->total.energy_evals++;
  int it = 0;
  for (; it < P.max_newton; ++it) {
    int nact = 0;
    for (int m = 0; m < M; ++m) {
      if (!act[m]) continue;
      if (!std::isfinite(E[m]) || !std::isfinite(gg[m]))
        fail(SSIM_ERR_NUMERIC, "mesh " + std::to_string(m) + ": non-finite energy or gradient");
      if (std::sqrt(gg[m]) < b->tol[m]) {
        act[m] = 0;
        b->stats[m].converged = 1;
        b->stats[m].newton_iterations = it;
      } else {
        ++nact;
      }
    }
    if (nact == 0) break;
    upload_mask(b, b->mask_newton, act);
    bassemble(b, b->mask_newton, u, 1);
    bpcg(b, b->mask_newton, g, -1.0, c->dir.p, iters);
    int maxk = 0;
    for (int m = 0; m < M; ++m)
      if (act[m]) {
        b->stats[m].hessian_evals++;
        b->stats[m].pcg_iterations += iters[m];
        maxk = std::max(maxk, iters[m]);
      }
    b->total.hessian_evals++;
    b->total.pcg_iterations += maxk;
    launch_bdot(b->view(b->mask_newton), b->nvwg, g, c->dir.p, c->part_d(), b->ev.p, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(b->pin_ev(), b->ev.p, sizeof(EvalScalars) * M, hipMemcpyDeviceToHost, s));
    sync(c);
    for (int m = 0; m < M; ++m) {
      gd[m] = act[m] ? b->pin_ev()[m].gd : 0.0;
      alpha[m] = 1.0;
      search[m] = act[m];
    }
    for (int ls = 0; ls < P.ls_max; ++ls) {
      upload_mask(b, b->mask_ls, search);
      HIPCHK(hipMemcpyAsync(b->alpha.p, alpha.data(), sizeof(double) * M, hipMemcpyHostToDevice, s));
      launch_baxpy_to(b->view(b->mask_ls), b->nvwg, u, b->alpha.p, c->dir.p, c->utry.p, s);
      beval(b, b->mask_ls, c->utry.p, c->gtry.p);
      b->total.energy_evals++;
      int left = 0;
      for (int m = 0; m < M; ++m) {
        if (!search[m]) continue;
        Et[m] = b->pin_ev()[m].energy;
        ggt[m] = b->pin_ev()[m].gg;
        b->stats[m].energy_evals++;
        if (Et[m] <= E[m] + P.ls_c * alpha[m] * gd[m]) {
          search[m] = 0;
        } else {
          alpha[m] *= P.ls_beta;
          ++left;
        }
      }
      if (!left) break;
    }
    // As in the single-mesh solver: a mesh whose line search exhausted its trials takes the last trial if that
    // did not raise its energy; otherwise it keeps its iterate and leaves the Newton loop unconverged.
    bool any_rejected = false;
    for (int m = 0; m < M; ++m) {
      search[m] = act[m] && search[m];
      if (!act[m]) continue;
      if (search[m]) {
        b->stats[m].ls_failures++;
        if (!(Et[m] <= E[m])) {
          b->stats[m].ls_rejected = 1;
          b->stats[m].newton_iterations = it + 1;
          act[m] = 0;
          any_rejected = true;
          continue;
        }
      }
      b->stats[m].last_change = std::fabs(E[m] - Et[m]);
      E[m] = Et[m];
      gg[m] = ggt[m];
    }
    if (any_rejected) upload_mask(b, b->mask_ls, act);  // copy back only the meshes that took their trial
    launch_bcopy2(b->view(any_rejected ? b->mask_ls : b->mask_newton), b->nvwg, c->utry.p, u, c->gtry.p, g, s);
    HIPCHK(hipGetLastError());
  }
  for (int m = 0; m < M; ++m)
    if (act[m]) b->stats[m].newton_iterations = it;  // stopped at the iteration cap
  b->total.newton_iterations = it;
  // polish: full Newton steps with the unprojected Hessian for every mesh, no line search
  for (int k = 0; k < P.newton_polish; ++k) {
    bassemble(b, b->mask_all, u, 0);
    bpcg(b, b->mask_all, g, -1.0, c->dir.p, iters);
    int maxk = 0;
    for (int m = 0; m < M; ++m) {
      b->stats[m].pcg_iterations += iters[m];
      maxk = std::max(maxk, iters[m]);
    }
    b->total.pcg_iterations += maxk;
    for (int m = 0; m < M; ++m) alpha[m] = 1.0;
    HIPCHK(hipMemcpyAsync(b->alpha.p, alpha.data(), sizeof(double) * M, hipMemcpyHostToDevice, s));
    launch_baxpy_to(b->view(b->mask_all), b->nvwg, u, b->alpha.p, c->dir.p, c->utry.p, s);
    HIPCHK(hipMemcpyAsync(u, c->utry.p, sizeof(double) * c->n, hipMemcpyDeviceToDevice, s));
    beval(b, b->mask_all, u, g);
    b->total.energy_evals++;
  }
}

}  // namespace

// ======================================================================================================
// extern "C" ABI

extern "C" {

ssim_status_t ssim_batch_create(ssim_batch** out, int nmesh, const int* nv, const int* nt, const double* vertices,
                                const int* cells, const double* lambda, const double* mu, double density,
                                double time_step) {
  if (!out) {
    g_err = "out is null";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  ssim_batch* b = nullptr;
  const ssim_status_t st = run([&] {
    if (nmesh <= 0) fail(SSIM_ERR_INVALID_ARGUMENT, "a batch needs at least one mesh");
    if (!nv || !nt || !vertices || !cells || !lambda || !mu) fail(SSIM_ERR_INVALID_ARGUMENT, "null batch arrays");
    if (!(time_step > 0)) fail(SSIM_ERR_INVALID_ARGUMENT, "Time step must be positive.");
    if (!(density >= 0)) fail(SSIM_ERR_INVALID_ARGUMENT, "Density must be positive, got " + std::to_string(density));
    long long pv = 0, pt = 0;
    for (int m = 0; m < nmesh; ++m) {
      if (nv[m] <= 0 || nt[m] <= 0) fail(SSIM_ERR_INVALID_ARGUMENT, "mesh " + std::to_string(m) + " is empty");
      pv += (nv[m] + kBatchRows - 1) / kBatchRows * (long long)kBatchRows;
      pt += (nt[m] + kBatchRows - 1) / kBatchRows * (long long)kBatchRows;
    }
    if (3 * pv >= (1LL << 31) || 12 * pt >= (1LL << 31)) fail(SSIM_ERR_INVALID_ARGUMENT, "batch too large for 32-bit indices");
    {
      long long off = 0, toff = 0;
      for (int m = 0; m < nmesh; ++m) {
        for (long long k = 0; k < 4LL * nt[m]; ++k) {
          const int v = cells[toff + k];
          if (v < 0 || v >= nv[m])
            fail(SSIM_ERR_INVALID_ARGUMENT, "mesh " + std::to_string(m) + ": cell " + std::to_string(k / 4) +
                                                " references vertex " + std::to_string(v) + " outside [0, " +
                                                std::to_string(nv[m]) + ")");
        }
        for (long long e = 0; e < nt[m]; ++e)
          if (!(lambda[off + e] > 0) || !(mu[off + e] > 0))
            fail(SSIM_ERR_INVALID_ARGUMENT, "mesh " + std::to_string(m) + ": Lamé parameters must be positive");
        off += nt[m];
        toff += 4LL * nt[m];
      }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) fail(SSIM_ERR_HIP, "no HIP device available");
    b = new ssim_batch();
    ssim_ctx* c = &b->c;
    b->nmesh = nmesh;
    b->nv_m.assign(nv, nv + nmesh);
    b->nt_m.assign(nt, nt + nmesh);
    b->vbase.resize(nmesh);
    b->tbase.resize(nmesh);
    b->hvwg_off.assign(nmesh + 1, 0);
    b->htwg_off.assign(nmesh + 1, 0);
    for (int m = 0; m < nmesh; ++m) {
      b->hvwg_off[m + 1] = b->hvwg_off[m] + (nv[m] + kBatchRows - 1) / kBatchRows;
      b->htwg_off[m + 1] = b->htwg_off[m] + (nt[m] + kBatchRows - 1) / kBatchRows;
      b->vbase[m] = kBatchRows * b->hvwg_off[m];
      b->tbase[m] = kBatchRows * b->htwg_off[m];
    }
    b->nvwg = b->hvwg_off[nmesh];
    b->ntwg = b->htwg_off[nmesh];
    c->mesh_rows.assign(b->vbase.begin(), b->vbase.end());
    c->mesh_rows.push_back((int)pv);
    c->nv = (int)pv;
    c->nt = (int)pt;
    c->n = 3 * c->nv;
    c->dt = time_step;
    c->density = density;
    c->per_tet = true;
    default_params(c->prm);
    c->hX.assign(3 * (size_t)pv, 0.0);
    c->hcells.assign(4 * (size_t)pt, -1);
    c->hlam.assign(pt, 1.0);
    c->hmu.assign(pt, 1.0);
    std::vector<int> vwg_mesh(b->nvwg), twg_mesh(b->ntwg), tend(nmesh), hvcoff(nmesh), htcoff(nmesh);
    {
      long long voff = 0, toff = 0;
      for (int m = 0; m < nmesh; ++m) {
        std::copy(vertices + 3 * voff, vertices + 3 * (voff + nv[m]), c->hX.begin() + 3 * (size_t)b->vbase[m]);
        for (long long k = 0; k < 4LL * nt[m]; ++k) c->hcells[4 * (size_t)b->tbase[m] + k] = cells[4 * toff + k] + b->vbase[m];
        std::copy(lambda + toff, lambda + toff + nt[m], c->hlam.begin() + b->tbase[m]);
        std::copy(mu + toff, mu + toff + nt[m], c->hmu.begin() + b->tbase[m]);
        for (int w = b->hvwg_off[m]; w < b->hvwg_off[m + 1]; ++w) vwg_mesh[w] = m;
        for (int w = b->htwg_off[m]; w < b->htwg_off[m + 1]; ++w) twg_mesh[w] = m;
        tend[m] = b->tbase[m] + nt[m];
        hvcoff[m] = (int)voff;
        htcoff[m] = (int)toff;
        voff += nv[m];
        toff += nt[m];
      }
      b->nv_compact = voff;
      b->nt_compact = toff;
    }
    hipStream_t s = c->stream;
    b->pin_bytes = (sizeof(EvalScalars) + 2 * sizeof(PcgScalars)) * (size_t)nmesh;
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&b->pin), b->pin_bytes, hipHostMallocDefault));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->pin), 64 * sizeof(double), hipHostMallocDefault));
    c->X.upload(c->hX, s);
    c->cells.upload(reinterpret_cast<const int4*>(c->hcells.data()), pt, s);
    b->vwg_mesh.upload(vwg_mesh, s);
    b->twg_mesh.upload(twg_mesh, s);
    b->vwg_off.upload(b->hvwg_off, s);
    b->twg_off.upload(b->htwg_off, s);
    b->tend.upload(tend, s);
    b->vcoff.upload(hvcoff, s);
    b->nvm.upload(b->nv_m, s);
    b->tcoff.upload(htcoff, s);
    std::vector<int> ones(nmesh, 1);
    b->mask_all.upload(ones, s);
    b->mask_newton.upload(ones, s);
    b->mask_ls.upload(ones, s);
    b->stats.assign(nmesh, ssim_step_stats{});
    batch_reset(b);
    for (int q = 0; q < 2; ++q) {
      b->slist[q].alloc(std::max(c->nslices, 1));
      b->wlist[q].alloc(std::max(b->nvwg, 1));
      HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&b->pin_list[q]),
                           sizeof(int4) * ((size_t)c->nslices + b->nvwg + 1), hipHostMallocDefault));
    }
  });
  if (st != SSIM_OK) {
    delete b;
    return st;
  }
  *out = b;
  return SSIM_OK;
}

ssim_status_t ssim_batch_destroy(ssim_batch* b) {
  if (!b) return SSIM_OK;
  (void)hipStreamSynchronize(b->c.stream);
  delete b;
  return SSIM_OK;
}

ssim_status_t ssim_batch_set_stream(ssim_batch* b, void* hip_stream) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    sync(&b->c);
    b->c.stream = reinterpret_cast<hipStream_t>(hip_stream);
  });
}

ssim_status_t ssim_batch_num_meshes(const ssim_batch* b, int* nmesh) {
  if (bguard(b) || !nmesh) return SSIM_ERR_INVALID_ARGUMENT;
  *nmesh = b->nmesh;
  return SSIM_OK;
}

ssim_status_t ssim_batch_set_threshold(ssim_batch* b, double rtol) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (!(rtol > 0)) fail(SSIM_ERR_INVALID_ARGUMENT, "threshold must be positive");
    b->c.rtol = rtol;
    for (int m = 0; m < b->nmesh; ++m) b->tol[m] = rtol * b->mass[m];
  });
}

ssim_status_t ssim_batch_set_solver_params(ssim_batch* b, const ssim_solver_params* p) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (!p) fail(SSIM_ERR_INVALID_ARGUMENT, "params is null");
    if (p->max_newton < 0 || !(p->pcg_rtol > 0) || p->pcg_maxit < 0 || !(p->ls_c > 0 && p->ls_c < 1) ||
        !(p->ls_beta > 0 && p->ls_beta < 1) || p->ls_max < 1 || p->newton_polish < 0 || p->newton_polish > 16)
      fail(SSIM_ERR_INVALID_ARGUMENT, "solver parameters out of range");
    if (p->pcg_precond != 0) fail(SSIM_ERR_INVALID_ARGUMENT, "batches solve with Jacobi-PCG only (pcg_precond = 0)");
    b->c.prm = *p;
  });
}

ssim_status_t ssim_batch_get_solver_params(const ssim_batch* b, ssim_solver_params* p) {
  if (bguard(b) || !p) return SSIM_ERR_INVALID_ARGUMENT;
  *p = b->c.prm;
  return SSIM_OK;
}

ssim_status_t ssim_batch_mark_dirichlet(ssim_batch* b, int mesh, const int* verts, int count, const double* targets) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    check_mesh(b, mesh, false);
    if (count < 0 || (count > 0 && !verts)) fail(SSIM_ERR_INVALID_ARGUMENT, "bad vertex list");
    std::vector<int> gv(count);
    for (int i = 0; i < count; ++i) {
      if (verts[i] < 0 || verts[i] >= b->nv_m[mesh]) fail(SSIM_ERR_INVALID_ARGUMENT, "vertex index out of range");
      gv[i] = verts[i] + b->vbase[mesh];
    }
    edit_dirichlet(&b->c, gv.data(), count, targets, true);
  });
}

ssim_status_t ssim_batch_mark_general(ssim_batch* b, int mesh, const int* verts, int count) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    check_mesh(b, mesh, false);
    if (count < 0 || (count > 0 && !verts)) fail(SSIM_ERR_INVALID_ARGUMENT, "bad vertex list");
    std::vector<int> gv(count);
    for (int i = 0; i < count; ++i) {
      if (verts[i] < 0 || verts[i] >= b->nv_m[mesh]) fail(SSIM_ERR_INVALID_ARGUMENT, "vertex index out of range");
      gv[i] = verts[i] + b->vbase[mesh];
    }
    edit_dirichlet(&b->c, gv.data(), count, nullptr, false);
  });
}

ssim_status_t ssim_batch_reset(ssim_batch* b) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] { batch_reset(b); });
}

ssim_status_t ssim_batch_prepare_step(ssim_batch* b) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    ssim_ctx* c = &b->c;
    launch_prepare(c->n, c->a(SSIM_ARRAY_DEFORMATION), c->a(SSIM_ARRAY_VELOCITY), c->a(SSIM_ARRAY_EXT_ACCEL), c->dof.p,
                   c->a(SSIM_ARRAY_DBC_VALUES), c->dt, c->a(SSIM_ARRAY_INERTIA), c->a(SSIM_ARRAY_NEXT_DEFORM),
                   c->stream);
    HIPCHK(hipGetLastError());
  });
}

ssim_status_t ssim_batch_compute_step(ssim_batch* b) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    ssim_ctx* c = &b->c;
    bnewton(b);
    // compute_step's final energy / gradient evaluation (time_step.hpp:520-528), for every mesh
    beval(b, b->mask_all, c->a(SSIM_ARRAY_NEXT_DEFORM), c->a(SSIM_ARRAY_FORCES));
    b->total.energy_evals++;
    int conv = 1;
    double gmax = 0, esum = 0;
    for (int m = 0; m < b->nmesh; ++m) {
      const EvalScalars& e = b->pin_ev()[m];
      if (!std::isfinite(e.energy) || !std::isfinite(e.gg))
        fail(SSIM_ERR_NUMERIC, "mesh " + std::to_string(m) + ": non-finite energy or gradient");
      b->stats[m].grad_norm = std::sqrt(e.gg);
      b->stats[m].energy = e.energy;
      b->stats[m].energy_evals++;
      conv &= b->stats[m].converged;
      gmax = std::max(gmax, b->stats[m].grad_norm);
      esum += e.energy;
    }
    b->total.converged = conv;
    b->total.grad_norm = gmax;
    b->total.energy = esum;
  });
}

ssim_status_t ssim_batch_step_next(ssim_batch* b) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    ssim_ctx* c = &b->c;
    launch_step_next(c->n, c->a(SSIM_ARRAY_NEXT_DEFORM), c->a(SSIM_ARRAY_DEFORMATION), c->a(SSIM_ARRAY_PREV_DEFORM),
                     c->a(SSIM_ARRAY_VELOCITY), c->dt, c->damping, c->stream);
    HIPCHK(hipGetLastError());
  });
}

ssim_status_t ssim_batch_step(ssim_batch* b) {
  ssim_status_t st = ssim_batch_prepare_step(b);
  if (st) return st;
  st = ssim_batch_compute_step(b);
  if (st) return st;
  return ssim_batch_step_next(b);
}

ssim_status_t ssim_batch_get_step_stats(const ssim_batch* b, int mesh, ssim_step_stats* stats) {
  if (bguard(b) || !stats) return SSIM_ERR_INVALID_ARGUMENT;
  if (mesh < -1 || mesh >= b->nmesh) {
    g_err = "mesh index out of range";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  *stats = mesh < 0 ? b->total : b->stats[mesh];
  return SSIM_OK;
}

ssim_status_t ssim_batch_get_array(ssim_batch* b, ssim_array_t which, int mesh, double* host_out) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (which < 0 || which >= SSIM_ARRAY_COUNT || !host_out) fail(SSIM_ERR_INVALID_ARGUMENT, "bad array request");
    check_mesh(b, mesh, true);
    ssim_ctx* c = &b->c;
    size_t off = 0;
    for (int m = 0; m < b->nmesh; ++m) {
      if (mesh >= 0 && m != mesh) continue;
      const size_t cnt = 3 * (size_t)b->nv_m[m];
      HIPCHK(hipMemcpyAsync(host_out + off, c->a(which) + 3 * (size_t)b->vbase[m], sizeof(double) * cnt,
                            hipMemcpyDeviceToHost, c->stream));
      off += cnt;
    }
    sync(c);
  });
}

ssim_status_t ssim_batch_set_array(ssim_batch* b, ssim_array_t which, int mesh, const double* host_in) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (which < 0 || which >= SSIM_ARRAY_COUNT || !host_in) fail(SSIM_ERR_INVALID_ARGUMENT, "bad array request");
    check_mesh(b, mesh, true);
    ssim_ctx* c = &b->c;
    size_t off = 0;
    for (int m = 0; m < b->nmesh; ++m) {
      if (mesh >= 0 && m != mesh) continue;
      const size_t cnt = 3 * (size_t)b->nv_m[m];
      HIPCHK(hipMemcpyAsync(c->a(which) + 3 * (size_t)b->vbase[m], host_in + off, sizeof(double) * cnt,
                            hipMemcpyHostToDevice, c->stream));
      off += cnt;
    }
    sync(c);
  });
}

ssim_status_t ssim_batch_get_array_device(ssim_batch* b, ssim_array_t which, double* dst) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (which < 0 || which >= SSIM_ARRAY_COUNT || !dst) fail(SSIM_ERR_INVALID_ARGUMENT, "bad array request");
    ssim_ctx* c = &b->c;
    launch_bunpad_vec(b->view(b->mask_all), b->nvwg, b->vcoff.p, b->nvm.p, c->a(which), dst, c->stream);
    HIPCHK(hipGetLastError());
    sync(c);
  });
}

ssim_status_t ssim_batch_set_array_device(ssim_batch* b, ssim_array_t which, const double* src) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (which < 0 || which >= SSIM_ARRAY_COUNT || !src) fail(SSIM_ERR_INVALID_ARGUMENT, "bad array request");
    ssim_ctx* c = &b->c;
    launch_bpad_vec(b->view(b->mask_all), b->nvwg, b->vcoff.p, b->nvm.p, src, c->a(which), c->stream);
    HIPCHK(hipGetLastError());
    sync(c);
  });
}

ssim_status_t ssim_batch_set_material(ssim_batch* b, const double* lambda, const double* mu) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (!lambda || !mu) fail(SSIM_ERR_INVALID_ARGUMENT, "null material arrays");
    ssim_ctx* c = &b->c;
    long long off = 0;
    for (int m = 0; m < b->nmesh; ++m) {
      for (int e = 0; e < b->nt_m[m]; ++e) {
        const double l = lambda[off + e], u = mu[off + e];
        if (!(l > 0) || !(u > 0) || !std::isfinite(l) || !std::isfinite(u))
          fail(SSIM_ERR_INVALID_ARGUMENT, "mesh " + std::to_string(m) + ": Lamé parameters must be positive and finite");
        c->hlam[b->tbase[m] + e] = l;
        c->hmu[b->tbase[m] + e] = u;
      }
      off += b->nt_m[m];
    }
    upload_material(c);  // also decides whether the material is uniform
  });
}

ssim_status_t ssim_batch_set_material_model(ssim_batch* b, ssim_material_t model) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  if (model != SSIM_MATERIAL_STABLE_NEOHOOKEAN && model != SSIM_MATERIAL_LINEAR) {
    g_err = "unknown material model";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  b->c.model = model;
  return SSIM_OK;
}

// Adjoint of one implicit step of every mesh at its converged u* = next_deform (see ssim_adjoint_step);
// device arrays in the compact layout
ssim_status_t ssim_batch_adjoint_step(ssim_batch* b, const double* grad_u, double* grad_uhat, double* grad_lambda,
                                      double* grad_mu, double* adjoint_out, int* pcg_iters_out) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return run([&] {
    if (!grad_u || !grad_uhat) fail(SSIM_ERR_INVALID_ARGUMENT, "grad_u and grad_uhat are required");
    ssim_ctx* c = &b->c;
    hipStream_t s = c->stream;
    const double dt2 = c->dt * c->dt;
    const double* u = c->a(SSIM_ARRAY_NEXT_DEFORM);
    const BatchView bt = b->view(b->mask_all);
    bassemble(b, b->mask_all, u, 0);
    launch_bpad_vec(bt, b->nvwg, b->vcoff.p, b->nvm.p, grad_u, c->utry.p, s);
    launch_mask_dbc(c->n, c->dof.p, c->utry.p, c->tmp.p, s);
    HIPCHK(hipGetLastError());
    std::vector<int> iters(b->nmesh, 0);
    bpcg(b, b->mask_all, c->tmp.p, 1.0, c->dir.p, iters);
    c->adj_lam.alloc(std::max(c->nt, 1));
    c->adj_mu.alloc(std::max(c->nt, 1));
    launch_belem_adjoint(c->mesh(), bt, b->ntwg, u, c->dir.p, dt2, c->gperm.p, c->gcontrib.p, c->adj_lam.p,
                         c->adj_mu.p, s);
    launch_bgrad_reduce(bt, b->nvwg, c->gseg.p, c->gcontrib.p, c->dof.p, c->gtry.p, c->part_g(), s);
    launch_bunpad_vec(bt, b->nvwg, b->vcoff.p, b->nvm.p, c->gtry.p, grad_uhat, s);
    if (grad_lambda) launch_bunpad_tet(bt, b->ntwg, b->tcoff.p, c->adj_lam.p, grad_lambda, s);
    if (grad_mu) launch_bunpad_tet(bt, b->ntwg, b->tcoff.p, c->adj_mu.p, grad_mu, s);
    if (adjoint_out) launch_bunpad_vec(bt, b->nvwg, b->vcoff.p, b->nvm.p, c->dir.p, adjoint_out, s);
    HIPCHK(hipGetLastError());
    sync(c);
    if (pcg_iters_out) std::copy(iters.begin(), iters.end(), pcg_iters_out);
  });
}

ssim_status_t ssim_batch_total_mass(const ssim_batch* b, int mesh, double* mass) {
  if (bguard(b) || !mass || mesh < 0 || mesh >= b->nmesh) return SSIM_ERR_INVALID_ARGUMENT;
  *mass = b->mass[mesh];
  return SSIM_OK;
}

ssim_status_t ssim_batch_get_layout_info(const ssim_batch* b, ssim_layout_info* info) {
  if (bguard(b) || !info) return SSIM_ERR_INVALID_ARGUMENT;
  return ssim_get_layout_info(&b->c, info);
}

ssim_status_t ssim_batch_set_tuning(ssim_batch* b, ssim_tune_t key, int value) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  if (key == SSIM_TUNE_SPMV_LANES && value == 4) {
    g_err = "batched PCG runs one row per lane (SpMV variant 0 or 1)";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  if (key == SSIM_TUNE_GRAD_TILE && value != kBlock) {
    g_err = "batches evaluate the gradient in 256-tet tiles (their per-mesh tet ranges are padded to 256)";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  if (key == SSIM_TUNE_COMPACT_PCT) {
    if (value < 0 || value > 100) {
      g_err = "compaction threshold must be a percentage in [0, 100]";
      return SSIM_ERR_INVALID_ARGUMENT;
    }
    b->compact_pct = value;
    return SSIM_OK;
  }
  if (key == SSIM_TUNE_PCG_GRAPH) {
    b->graph_mode = value != 0;
    return SSIM_OK;
  }
  b->drop_graphs();  // launch shapes may change
  return ssim_set_tuning(&b->c, key, value);
}

ssim_status_t ssim_batch_get_tuning(const ssim_batch* b, ssim_tune_t key, int* value) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  if (key == SSIM_TUNE_COMPACT_PCT || key == SSIM_TUNE_PCG_GRAPH) {
    if (!value) return SSIM_ERR_INVALID_ARGUMENT;
    *value = key == SSIM_TUNE_COMPACT_PCT ? b->compact_pct : b->graph_mode;
    return SSIM_OK;
  }
  return ssim_get_tuning(&b->c, key, value);
}

ssim_status_t ssim_batch_set_kernel_timing(ssim_batch* b, int enabled) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return ssim_set_kernel_timing(&b->c, enabled);
}

ssim_status_t ssim_batch_get_kernel_timing(const ssim_batch* b, ssim_kernel_t kernel, double* avg_ms,
                                           long long* samples) {
  if (bguard(b)) return SSIM_ERR_INVALID_ARGUMENT;
  return ssim_get_kernel_timing(&b->c, kernel, avg_ms, samples);
}

}  // extern "C"
