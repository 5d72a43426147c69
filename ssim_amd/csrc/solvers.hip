// solvers.hip -- the host drivers of the steppers: energy and gradient evaluation, Hessian assembly, Jacobi-PCG
// in chunks of device-resident iterations, Newton with backtracking (stepper/backward_euler.hpp:9-41), L-BFGS in
// the Gram and the vector-by-vector form with the diagonal, projective-dynamics and Hessian-at-start
// preconditioners (backward_euler.hpp:43-67, stepper/projective_dynamics.hpp:9-150), nonlinear CG with an
// optional external preconditioner (stepper/nonlinear_cg.hpp:9-97) and forward Euler (time_step.hpp:717-736).
// The host decides control flow only; the arithmetic runs in the kernels of kernels.hip.
#include "context.h"

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace ssim_hip;

namespace ssim_hip {

// update_energy_and_gradients at deformation u → gradient into g; returns E and ‖g‖²
EvalOut eval_energy_grad(ssim_ctx* c, const double* u, double* g) {
  hipStream_t s = c->stream;
  const double dt2 = c->dt * c->dt;
  int ne = 0, ng = 0;
  timed_ext(c, SSIM_KERNEL_ELEM_GRAD, [&](hipEvent_t e0, hipEvent_t e1) {
    ne = launch_elem_grad_tile(c->mesh(), c->grad_tile, u, c->a(SSIM_ARRAY_INERTIA), dt2, c->gt_slot.p, c->gt_off.p,
                               c->gt_run.p, c->gpart.p, c->part_e(), c->prm.fp32_elements != 0, s, e0, e1);
  });
  timed_ext(c, SSIM_KERNEL_GRAD_REDUCE, [&](hipEvent_t e0, hipEvent_t e1) {
    ng = launch_grad_gather(c->nv, c->gv_seg.p, c->gv_pos.p, c->gpart.p, c->dof.p, g, c->part_g(), s, e0, e1);
  });
  launch_finalize(c->part_e(), ne, 1, c->part_g(), ng, &c->ev.p->energy, s);  // energy, gg
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->pin, c->ev.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  sync(c);
  c->stats.energy_evals++;
  return EvalOut{c->pin[0], c->pin[1]};
}

static double device_dot(ssim_ctx* c, const double* a, const double* b) {
  launch_dot(c->n, a, b, c->part_d(), &c->ev.p->dot, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->pin, &c->ev.p->dot, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  sync(c);
  return c->pin[0];
}


void ensure_hcontrib(ssim_ctx* c) {
  if (!c->hcontrib.p) c->hcontrib.alloc(90 * (size_t)std::max(c->nt, 1));
}

// Hessian assembly at deformation u (dt²-scaled stiffness plus the consistent mass) into the blocked-ELL values,
// the Jacobi inverse diagonal dinv and, when diag is not null, the diagonal: the tile-reduced kernels, or with
// hess_tile = 0 the per-tet contribution slots and their per-block reduce
void assemble_into(ssim_ctx* c, const double* u, double dt2, int make_spsd, int filter_dbc, double* values,
                   double* dinv, double* diag, bool fp32) {
  hipStream_t s = c->stream;
  if (c->hess_tile) {
    timed_ext(c, SSIM_KERNEL_ELEM_HESS, [&](hipEvent_t e0, hipEvent_t e1) {
      launch_elem_hess_tile(c->mesh(), u, dt2, make_spsd, c->ht_slot.p, c->ht_off.p, c->ht_run.p, c->hpart.p,
                            c->hparts, fp32, s, e0, e1);
    });
    timed_ext(c, SSIM_KERNEL_HESS_REDUCE, [&](hipEvent_t e0, hipEvent_t e1) {
      launch_hess_gather(c->nub, c->hb_seg.p, c->hb_pos.p, c->hpart.p, c->hparts, c->ub_rc.p, c->ub_pos.p, c->dof.p,
                         filter_dbc, values, dinv, diag, s, e0, e1);
    });
  } else {
    ensure_hcontrib(c);
    timed(c, SSIM_KERNEL_ELEM_HESS, [&] {
      launch_elem_hess(c->mesh(), u, dt2, make_spsd, c->hperm.p, c->hcontrib.p, fp32, s);
    });
    timed(c, SSIM_KERNEL_HESS_REDUCE, [&] {
      launch_hess_reduce(c->nub, c->hseg.p, c->hcontrib.p, c->ub_rc.p, c->ub_pos.p, c->dof.p, filter_dbc, values,
                         dinv, diag, s);
    });
  }
  HIPCHK(hipGetLastError());
}

// update_hessian at deformation u into the blocked-ELL values and the Jacobi diagonal
void assemble_hessian(ssim_ctx* c, const double* u, int make_spsd, int filter_dbc, double dt2,
                                double* values) {
  assemble_into(c, u, dt2, make_spsd, filter_dbc, values, c->dinv.p, nullptr, c->prm.fp32_elements != 0);
  c->stats.hessian_evals++;
}


void create_poll_events(ssim_ctx* c) {
  for (auto& e : c->poll_ev)
    if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

// The context's own PCG work buffers.
PcgWork ctx_pcg_work(ssim_ctx* c) {
  return PcgWork{c->r.p, c->z.p, c->p.p, c->q.p, c->part_pa(), c->part_pb(0), c->part_pb(1), c->pcg.p, c->pcg_it.p};
}

// Enqueues `count` PCG iterations (kernels A and B) on stream s.
static void enqueue_pcg_iterations(ssim_ctx* c, const PcgWork& w, const BellView& A, const double* dinv, double* x,
                                   int nb, int na, int count, hipStream_t s) {
  for (int j = 0; j < count; ++j) {
    launch_pcg_spmv(A, w.z, w.p, w.q, w.pb0, w.pb1, nb, w.it, w.sc, w.pa, x, s);
    launch_pcg_update(c->n, dinv, w.r, w.z, w.q, w.pa, na, w.pb0, w.pb1, w.it, w.sc, s);
  }
}

// The graph of `count` iterations on matrix A writing x with the context's work buffers, captured on first use
// (buffers are fixed until the next reset, which drops the cache).
static hipGraphExec_t pcg_chunk_graph(ssim_ctx* c, const BellView& A, const double* dinv, double* x, int nb, int na,
                               int count) {
  for (const auto& g : c->pcg_graphs)
    if (g.count == count && g.x == x && g.val == A.val) return g.exec;
  if (!c->cap_stream) HIPCHK(hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
  HIPCHK(hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal));
  hipGraph_t graph = nullptr;
  try {
    enqueue_pcg_iterations(c, ctx_pcg_work(c), A, dinv, x, nb, na, count, c->cap_stream);
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

// The preconditioner of plain CG (pcg_precond = 1): z = r, as a diagonal of ones in Jacobi's place.  Made when
// the parameter is set and at reset, never inside a solve.
void ensure_unit_diagonal(ssim_ctx* c) {
  if (c->ones.count != (size_t)std::max(c->n, 1)) {
    c->ones.alloc(std::max(c->n, 1));
    std::vector<double> h(c->ones.count, 1.0);
    HIPCHK(hipMemcpyAsync(c->ones.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->stream));
    sync(c);
  }
}
static const double* unit_diagonal(ssim_ctx* c) {
  if (!c->ones.p || c->ones.count != (size_t)std::max(c->n, 1)) fail(SSIM_ERR_STATE, "internal: plain-CG diagonal missing");
  return c->ones.p;
}

// Jacobi-PCG: x = A⁻¹ (sign·b) for the blocked-ELL values `val` with inverse diagonal `dinv` (default: the
// current Hessian).  Returns iterations.  `work` (default: the context's buffers) holds r, z, p, q, the partial
// sums and the PCG scalars; with a caller's work the chunks are always launched kernel by kernel (no graphs).
int pcg_solve(ssim_ctx* c, const double* b, double sign, double* x, double rtol, int maxit, const double* val,
              const double* dinv, const PcgWork* work) {
  hipStream_t s = c->stream;
  if (!val) val = c->val.p;
  if (!dinv) dinv = c->prm.pcg_precond == 1 ? unit_diagonal(c) : c->dinv.p;
  const PcgWork w = work ? *work : ctx_pcg_work(c);
  const bool graphs = c->pcg_graph && !work;
  const BellView A = c->bell(val);
  const int nb = launch_pcg_init(c->n, b, sign, dinv, x, w.r, w.z, rtol, maxit, w.sc, w.it, w.pb0, s);
  HIPCHK(hipGetLastError());
  create_poll_events(c);
  const int na = pcg_spmv_grid(A);
  // Chunks of iterations are enqueued one ahead of the done-flag poll, so the GPU never drains while the host
  // reads a flag; kernels enqueued after convergence return at entry.  The first iteration of a chunk is launched
  // on its own (it carries the sampled kernel timers); the others are replayed from a captured graph.
  // each poll reads rr, bb, tol2, the ρ slots and done, iters, maxit, breakdown: 64 bytes of PcgScalars from rr on
  struct Poll {
    double rr, bb, tol2, pad, rz_hist[2];
    int done, iters, maxit, breakdown;
  };
  static_assert(sizeof(Poll) == 64 && offsetof(PcgScalars, breakdown) - offsetof(PcgScalars, rr) ==
                                          offsetof(Poll, breakdown), "the poll mirrors PcgScalars from rr on");
  Poll* polls = reinterpret_cast<Poll*>(c->pin);  // two slots
  auto enqueue = [&](int count, int slot) {
    timed_ext(c, SSIM_KERNEL_PCG_SPMV, [&](hipEvent_t e0, hipEvent_t e1) {
      launch_pcg_spmv(A, w.z, w.p, w.q, w.pb0, w.pb1, nb, w.it, w.sc, w.pa, x, s, e0, e1);
    });
    timed_ext(c, SSIM_KERNEL_PCG_UPDATE, [&](hipEvent_t e0, hipEvent_t e1) {
      launch_pcg_update(c->n, dinv, w.r, w.z, w.q, w.pa, na, w.pb0, w.pb1, w.it, w.sc, s, e0, e1);
    });
    if (count > 1) {
      if (graphs)
        HIPCHK(hipGraphLaunch(pcg_chunk_graph(c, A, dinv, x, nb, na, count - 1), s));
      else
        enqueue_pcg_iterations(c, w, A, dinv, x, nb, na, count - 1, s);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(polls + slot, &w.sc->rr, sizeof(Poll), hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(c->poll_ev[slot], s));
  };
  // Chunk sizes double from 8 up to pcg_chunk_max while the solve is far from converged.  Near the end the
  // residual history sizes them: from the log-linear decrease of ‖r‖² between the last two polls, the chunk
  // after the one in flight covers the predicted remainder with a 10 % + 4 margin (at least 4 iterations, so
  // the GPU stays fed while the host polls).  Launches after the solve stopped return early, but each still
  // costs a few microseconds; this keeps their number to about the margin instead of up to two full chunks.
  const int cmax = std::max(8, c->pcg_chunk_max);
  int launched = 0, slot = 0, inflight = 8;
  enqueue(inflight, slot);
  launched += inflight;
  int next = std::min(16, cmax);
  int k_prev = -1;
  double rr_prev = 0.0;
  for (;;) {
    enqueue(next, slot ^ 1);
    launched += next;
    wait_event(c->poll_ev[slot]);
    if (c->timing) harvest_timers(c);
    const Poll& f = polls[slot];
    if (f.done) {
      if (f.breakdown == 1) fail(SSIM_ERR_NUMERIC, "PCG breakdown: pᵀAp ≤ 0 (matrix not positive definite)");
      if (f.breakdown == 2) fail(SSIM_ERR_NUMERIC, "PCG breakdown: NaN residual");
      return f.iters;
    }
    if (launched > maxit + 2 * cmax + 16) fail(SSIM_ERR_STATE, "internal: PCG did not stop");
    // f describes the state after the chunk that just completed (iteration f.iters); `next` is now in flight
    int want = std::min(next * 2, cmax);
    if (k_prev >= 0 && f.iters > k_prev && f.rr > 0.0 && rr_prev > 0.0 && f.tol2 > 0.0) {
      const double slope = std::log(f.rr / rr_prev) / (f.iters - k_prev);  // log ‖r‖² per iteration
      if (slope < 0.0) {
        const double remaining = std::log(f.tol2 / f.rr) / slope - next;  // after the chunk in flight
        if (remaining < want) want = std::max(4, (int)std::ceil(1.1 * remaining + 4.0));
      }
    }
    k_prev = f.iters;
    rr_prev = f.rr;
    next = std::min(want, cmax);
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
    if (!ok) {
      c->stats.ls_failures++;
      if (!(trial.E <= cur.E)) {  // even the smallest trial raises the energy: keep the iterate and stop
        c->stats.ls_rejected = 1;
        ++it;
        break;
      }
    }
    HIPCHK(hipMemcpyAsync(u, c->utry.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(g, c->gtry.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    c->stats.last_change = std::fabs(cur.E - trial.E);
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
static double pd_diffusivity(ssim_material_t model, double lam, double mu) {
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
  ensure_hcontrib(c);
  launch_elem_pd(c->mesh(), c->dt * c->dt * kappa, c->hperm.p, c->hcontrib.p, s);
  launch_hess_reduce(c->nub, c->hseg.p, c->hcontrib.p, c->ub_rc.p, c->ub_pos.p, c->dof.p, 1, c->pd_val.p,
                     c->pd_dinv.p, nullptr, s);
  HIPCHK(hipGetLastError());
  sync(c);
}

// H0 of the L-BFGS stepper: diagonal of the SPD-projected, Dirichlet-filtered Hessian at rest (u = 0)
static void ensure_hrest(ssim_ctx* c) {
  if (c->hrest.p) return;
  const int n = c->n;
  hipStream_t s = c->stream;
  c->hrest.alloc(n);
  HIPCHK(hipMemsetAsync(c->tmp.p, 0, sizeof(double) * n, s));
  const double dt2 = c->dt * c->dt;
  assemble_into(c, c->tmp.p, dt2, 1, 1, c->val.p, c->dinv.p, c->hrest.p, c->prm.fp32_elements != 0);
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
static void lbfgs_solve_gram(ssim_ctx* c) {
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
  // The iterate and its gradient alternate between the output arrays and the trial buffers: an accepted trial
  // becomes the iterate by exchanging the pointers, and the previous iterate stays in the trial buffers until the
  // next Gram pass has formed the new pair from both.  The result is moved to the output arrays at the end.
  double* const u_out = c->a(SSIM_ARRAY_NEXT_DEFORM);
  double* const g_out = c->a(SSIM_ARRAY_FORCES);
  double *u = u_out, *g = g_out, *ut = c->utry.p, *gt = c->gtry.p;
  double* d = c->dir.p;  // r ≈ H⁻¹g; the search direction is −r
  EvalOut cur = eval_energy_grad(c, u, g);
  int it = 0, hist = 0, head = 0, fresh = -1;  // fresh: slot of the pair of the last accepted step
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
      launch_lbfgs_gram(n, c->lb_S.p, c->lb_Y.p, g, c->hrest.p, fresh, u, ut, gt, c->lb_part.p, c->lb_scal.p, s);
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
      launch_lbfgs_dir(n, c->lb_S.p, c->lb_Y.p, g, c->hrest.p, cf, d, u, -1.0, ut, s);  // and ut = u − d
      HIPCHK(hipGetLastError());
      gd = -gr;
    }
    double alpha = 1.0;
    EvalOut trial{0, 0};
    bool ok = false;
    for (int ls = 0; ls < c->prm.lbfgs_ls_max; ++ls) {
      if (ls > 0 || vector_mode) launch_axpy_to(n, u, -alpha, d, ut, s);
      trial = eval_energy_grad(c, ut, gt);
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
    if (vector_mode) {  // the vector recursion reads the pair from its slot: store it and move the iterate
      launch_lbfgs_shift(n, ut, gt, u, g, c->lb_S.p + (size_t)head * n, c->lb_Y.p + (size_t)head * n, s);
      HIPCHK(hipGetLastError());
    } else {  // the next Gram pass forms and stores the pair
      std::swap(u, ut);
      std::swap(g, gt);
    }
    fresh = head;
    c->stats.last_change = std::fabs(cur.E - trial.E);
    cur = trial;
  }
  if (u != u_out) {
    HIPCHK(hipMemcpyAsync(u_out, u, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(g_out, g, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
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
void lbfgs_solve(ssim_ctx* c, bool pd, bool h0) {
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
    c->stats.last_change = std::fabs(cur.E - trial.E);
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
static void apply_precond(ssim_ctx* c, const double* g, double* a, double* out) {
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
    c->stats.last_change = std::fabs(cur.E - trial.E);
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

}  // namespace ssim_hip
