// context.hip -- host side of libssim_hip.so: errors, parameters, timers and the extern "C" ABI declared in
// include/ssim_hip.h (the setup done at reset is in topology.hip, the stepper drivers in solvers.hip).
//
// The host drives control flow only (Newton iterations, line-search decisions, PCG chunking); every
// floating-point operation of a step runs in the HIP kernels of kernels.hip.  PCG scalars stay on the device
// and the host polls a done-flag once per chunk of iterations.
#include "context.h"

#include <algorithm>
#include <chrono>
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
  p.pcg_precond = 0;
}

void ssim_hip::check_material(double youngs, double poisson, double density) {
  // time_step.hpp:369-374
  if (!(youngs >= 0)) fail(SSIM_ERR_INVALID_ARGUMENT, "Young's modulus must be positive, got " + std::to_string(youngs));
  if (!(poisson > 0 && poisson < 0.5))
    fail(SSIM_ERR_INVALID_ARGUMENT, "Poisson's ratio must be in (0, 0.5), got " + std::to_string(poisson));
  if (!(density >= 0)) fail(SSIM_ERR_INVALID_ARGUMENT, "Density must be positive, got " + std::to_string(density));
}

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

// Waits for the context's stream (sync) or for an event (wait_event).  The waits here are scalar read-backs and PCG
// polls a few microseconds to a few milliseconds long, so the host polls for up to 2 ms before it blocks: a blocking
// wait adds the thread's wake-up latency to every read-back (L-BFGS on the 1.57M-tet cube: 192 → 185-189 µs per
// iteration with polling).
namespace {
template <typename Query, typename Block>
hipError_t poll_then_block(Query&& query, Block&& block) {
  const auto t0 = std::chrono::steady_clock::now();
  hipError_t e;
  while ((e = query()) == hipErrorNotReady)
    if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) return block();
  return e;
}
}  // namespace

void ssim_hip::wait_event(hipEvent_t ev) {
  HIPCHK(poll_then_block([&] { return hipEventQuery(ev); }, [&] { return hipEventSynchronize(ev); }));
}

void ssim_hip::sync(ssim_ctx* c) {
  HIPCHK(poll_then_block([&] { return hipStreamQuery(c->stream); }, [&] { return hipStreamSynchronize(c->stream); }));
  if (c->timing) harvest_timers(c);
}

namespace {

ssim_status_t guard(ssim_ctx* c) {
  if (!c) {
    g_err = "null context";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  return SSIM_OK;
}

}  // namespace

void ssim_hip::require_reset(ssim_ctx* c) {
  if (!c->has_reset) fail(SSIM_ERR_STATE, "context not reset");
}

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
    HIPCHK(hipGetDevice(&c->device));
    noalloc_scratch_init();
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
    if (n == 0) return;
    launch_material_eval(dim, model, n, F, lambda, mu, make_spsd, psi, P, H, static_cast<hipStream_t>(stream));
    HIPCHK(hipGetLastError());
  });
}

ssim_status_t ssim_tri_eval(ssim_material_t model, int nv, int nt, const double* X, const int* cells, const double* u,
                            const double* lambda, const double* mu, double density, int make_spsd, double* area,
                            double* dminv, double* F, double* energy, double* grad, double* stiff, double* mass,
                            double* laplacian, void* stream) {
  return run([&] {
    if (model != SSIM_MATERIAL_STABLE_NEOHOOKEAN && model != SSIM_MATERIAL_LINEAR)
      fail(SSIM_ERR_INVALID_ARGUMENT, "unknown material model");
    if (nv < 0 || nt < 0) fail(SSIM_ERR_INVALID_ARGUMENT, "nv and nt must be non-negative");
    if (nt > 0 && (!X || !cells || !u || !lambda || !mu)) fail(SSIM_ERR_INVALID_ARGUMENT, "null input array");
    if (nt == 0) return;
    const TriOut o{area, dminv, F, energy, grad, stiff, mass, laplacian};
    launch_tri_eval(model, nv, nt, X, cells, u, lambda, mu, density, make_spsd, o, static_cast<hipStream_t>(stream));
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
        !(p->lbfgs_c > 0 && p->lbfgs_c < 1) || p->lbfgs_ls_max < 1 || p->newton_polish < 0 || p->newton_polish > 16 ||
        p->pcg_precond < 0 || p->pcg_precond > 1)
      fail(SSIM_ERR_INVALID_ARGUMENT, "solver parameter out of range");
    const bool fp32_changed = (p->fp32_elements != 0) != (ctx->prm.fp32_elements != 0);
    ctx->prm = *p;
    if (p->pcg_precond == 1) ensure_unit_diagonal(ctx);  // allocated here, so that no solve allocates
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

ssim_status_t ssim_device_dof_type(ssim_ctx* ctx, unsigned char** device_ptr) {
  if (guard(ctx)) return SSIM_ERR_INVALID_ARGUMENT;
  if (!device_ptr) {
    g_err = "device_ptr is null";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  *device_ptr = ctx->dof.p;
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
  if (!ctx || !nv) {
    g_err = "null argument";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  *nv = ctx->nv;
  return SSIM_OK;
}
ssim_status_t ssim_num_cells(const ssim_ctx* ctx, int* nt) {
  if (!ctx || !nt) {
    g_err = "null argument";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  *nt = ctx->nt;
  return SSIM_OK;
}
ssim_status_t ssim_total_mass(const ssim_ctx* ctx, double* mass) {
  if (!ctx || !mass) {
    g_err = "null argument";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  *mass = ctx->total_mass;
  return SSIM_OK;
}
ssim_status_t ssim_grad_threshold_abs(const ssim_ctx* ctx, double* tol) {
  if (!ctx || !tol) {
    g_err = "null argument";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
  *tol = ctx->tol_abs;
  return SSIM_OK;
}

ssim_status_t ssim_csr_nnz(const ssim_ctx* ctx, long long* nnz) {
  if (!ctx || !nnz) {
    g_err = "null argument";
    return SSIM_ERR_INVALID_ARGUMENT;
  }
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
    std::vector<double> hv(c->val.count);
    HIPCHK(hipMemcpyAsync(hv.data(), dval, sizeof(double) * hv.size(), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    long long p = 0;
    for (int v = 0; v < nv; ++v) {
      const int len = c->brow[v + 1] - c->brow[v];
      const int kd = c->kdiag[v];
      for (int b = 0; b < 3; ++b)
        for (int k = 0; k < len; ++k) {
          const int w = c->bcol[c->brow[v] + k];
          if (k >= kd) {
            const size_t bs = c->valpos(v, c->bslot[c->brow[v] + k]);
            const bool sym6 = c->fixed_layout && k == kd;
            for (int d = 0; d < 3; ++d) values[p++] = hv[bs + ssim_ctx::val_entry(sym6, b, d)];
          } else {  // lower block (v, w) = transpose of the stored (w, v)
            const int* bw = c->bcol.data() + c->brow[w];
            const int kw = (int)(std::lower_bound(bw, bw + (c->brow[w + 1] - c->brow[w]), v) - bw);
            const size_t bs = c->valpos(w, c->bslot[c->brow[w] + kw]);
            for (int d = 0; d < 3; ++d) values[p++] = hv[bs + ssim_ctx::val_entry(false, d, b)];
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
      assemble_into(ctx, ctx->a(SSIM_ARRAY_NEXT_DEFORM), 0.0, 0, filtered, ctx->massval.p, ctx->tmp.p, nullptr, false);
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
  info->fixed_layout = ctx->fixed_layout ? 1 : 0;
  info->reserved = 0;
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
        if (value != 0 && ctx->fixed_layout)
          fail(SSIM_ERR_INVALID_ARGUMENT, "this mesh has the fixed layout, which only the fixed-width SpMV (0) reads");
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
      case SSIM_TUNE_GRAD_TILE:
        if (value != 64 && value != 128 && value != kBlock)
          fail(SSIM_ERR_INVALID_ARGUMENT, "gradient tile must be 64, 128 or 256 tets");
        if (value != ctx->grad_tile) {
          ctx->grad_tile = value;
          if (ctx->nt > 0 && !ctx->hcells.empty()) build_grad_tiles(ctx);
        }
        break;
      case SSIM_TUNE_HESS_TILE:
        ctx->hess_tile = value != 0;
        break;
      case SSIM_TUNE_REST_FROM_X:
        ctx->rest_from_x = value != 0;
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
    case SSIM_TUNE_GRAD_TILE:
      *value = ctx->grad_tile;
      return SSIM_OK;
    case SSIM_TUNE_HESS_TILE:
      *value = ctx->hess_tile;
      return SSIM_OK;
    case SSIM_TUNE_REST_FROM_X:
      *value = ctx->rest_from_x;
      return SSIM_OK;
    default:
      g_err = "unknown tuning key";
      return SSIM_ERR_INVALID_ARGUMENT;
  }
}

}  // extern "C"
