// Host side of the differentiable time step (pehip_adjoint_*): checkpointed forward runs and the
// reverse pass.  The reverse pass re-runs each segment between checkpoints from its checkpoint,
// records per step the contacts in sweep order with their converged impulses, the positions at
// the start of the step, the solver body records and the post-solve velocities, and then walks
// the segment's steps backwards:
//   1. alpha = adjoint of the post-solve velocities (v' gets dt * xbar from x' = x + v' dt);
//   2. local Jacobians of every contact (dual numbers) and the per-contact 3x3 systems;
//   3. Gauss-Seidel sweeps, colour by colour in the forward's colouring, on the transposed
//      linearised fixed-point system at the converged impulses (implicit differentiation at
//      fixed active set: each contact keeps its mode -- separating, static, sliding);
//   4. xbar += geometric terms, mubar += friction terms, (vbar, wbar) = alpha.
#include <algorithm>
#include <cstring>
#include <vector>

#include "adjoint.hpp"
#include "context.hpp"


namespace {

constexpr int kStateDoubles = 17;  // pos 3, q 4, v 3, w 3, plane normal 3, plane d 1

struct StepTape {
  uint32_t nc = 0, cap = 0;
  std::vector<uint32_t> colour_off;
  uint32_t *b1 = nullptr, *b2 = nullptr;
  double *p = nullptr, *mu = nullptr, *pos = nullptr, *sb = nullptr, *vel = nullptr;
};

}  // namespace

struct AdjointState {
  double dt = 0;
  uint32_t nsteps = 0, every = 0;
  std::vector<double*> checkpoints;  // kStateDoubles * n each, at steps 0, every, 2*every, ...
  double* final_state = nullptr;
  pehip_params prm{};
  // reverse-pass work space
  std::vector<StepTape> tape;
  uint32_t ccap = 0;
  double *jac = nullptr, *minv = nullptr, *lambda = nullptr, *alpha = nullptr;
  double *xbar = nullptr, *vbar = nullptr, *wbar = nullptr, *mubar = nullptr;
  unsigned long long* delta = nullptr;
};

static void free_tape(StepTape& t) {
  dfree(t.b1); dfree(t.b2); dfree(t.p); dfree(t.mu); dfree(t.pos); dfree(t.sb); dfree(t.vel);
  t.cap = 0;
  t.nc = 0;
}

void pe_adjoint_free(pehip_ctx* ctx) {
  AdjointState* a = ctx->adj;
  if (!a) return;
  for (auto& c : a->checkpoints) dfree(c);
  dfree(a->final_state);
  for (auto& t : a->tape) free_tape(t);
  dfree(a->jac); dfree(a->minv); dfree(a->lambda); dfree(a->alpha);
  dfree(a->xbar); dfree(a->vbar); dfree(a->wbar); dfree(a->mubar); dfree(a->delta);
  delete a;
  ctx->adj = nullptr;
}

static void save_state(pehip_ctx* ctx, double* dst) {
  const size_t n = ctx->n, D = sizeof(double);
  const hipStream_t s = ctx->stream;
  hipMemcpyAsync(dst, ctx->pos, 3 * D * n, hipMemcpyDeviceToDevice, s);
  hipMemcpyAsync(dst + 3 * n, ctx->q, 4 * D * n, hipMemcpyDeviceToDevice, s);
  hipMemcpyAsync(dst + 7 * n, ctx->v, 3 * D * n, hipMemcpyDeviceToDevice, s);
  hipMemcpyAsync(dst + 10 * n, ctx->w, 3 * D * n, hipMemcpyDeviceToDevice, s);
  hipMemcpyAsync(dst + 13 * n, ctx->pnormal, 3 * D * n, hipMemcpyDeviceToDevice, s);
  hipMemcpyAsync(dst + 16 * n, ctx->pd, D * n, hipMemcpyDeviceToDevice, s);
}

static void restore_state(pehip_ctx* ctx, const double* src) {
  const size_t n = ctx->n, D = sizeof(double);
  const hipStream_t s = ctx->stream;
  hipMemcpyAsync(ctx->pos, src, 3 * D * n, hipMemcpyDeviceToDevice, s);
  hipMemcpyAsync(ctx->q, src + 3 * n, 4 * D * n, hipMemcpyDeviceToDevice, s);
  hipMemcpyAsync(ctx->v, src + 7 * n, 3 * D * n, hipMemcpyDeviceToDevice, s);
  hipMemcpyAsync(ctx->w, src + 10 * n, 3 * D * n, hipMemcpyDeviceToDevice, s);
  hipMemcpyAsync(ctx->pnormal, src + 13 * n, 3 * D * n, hipMemcpyDeviceToDevice, s);
  hipMemcpyAsync(ctx->pd, src + 16 * n, D * n, hipMemcpyDeviceToDevice, s);
  adj_restore_R((uint32_t)n, ctx->q, ctx->R, s);
  launch_aabb(ctx->bodies(), s);
}

static pehip_status check_supported(pehip_ctx* ctx, const pehip_params& P) {
  if (P.relaxation_model != PEHIP_INELASTIC_FRICTIONLESS &&
      P.relaxation_model != PEHIP_APPROX_INELASTIC_COULOMB_DECOUPLING &&
      P.relaxation_model != PEHIP_INELASTIC_COULOMB_DECOUPLING)
    return fail(ctx, PEHIP_NOT_SUPPORTED, "the adjoint supports relaxation models 0, 1 and 3");
  if (P.sweep_mode != PEHIP_SWEEP_COLOURED) return fail(ctx, PEHIP_NOT_SUPPORTED, "the adjoint runs in coloured mode");
  if (P.damping < 1.0) return fail(ctx, PEHIP_NOT_SUPPORTED, "the adjoint does not support damping");
  for (uint32_t i = 0; i < ctx->n; ++i)
    if (ctx->h_shape[i] != PEHIP_SPHERE && ctx->h_shape[i] != PEHIP_PLANE)
      return fail(ctx, PEHIP_NOT_SUPPORTED, "the adjoint supports spheres and planes");
  return PEHIP_OK;
}

// Records one step of the tape from the context right after pe_one_step.
static pehip_status record(pehip_ctx* ctx, AdjointState* a, StepTape& t, const double* pos_before) {
  const hipStream_t s = ctx->stream;
  const size_t n = ctx->n, D = sizeof(double);
  const uint32_t nc = ctx->nc;
  // each array is allocated on its own and a capacity is recorded only once its arrays all exist,
  // so a failed allocation is retried by the next call instead of leaving a null array in use
  if (!t.pos) CK(dalloc(t.pos, 3 * n));
  if (!t.sb) CK(dalloc(t.sb, (size_t)kSB * n));
  if (!t.vel) CK(dalloc(t.vel, 6 * n));
  if (nc > t.cap) {
    dfree(t.b1); dfree(t.b2); dfree(t.p); dfree(t.mu);
    t.cap = 0;
    const uint32_t cap = std::max<uint32_t>(nc + nc / 4, 1024);
    CK(dalloc(t.b1, cap)); CK(dalloc(t.b2, cap)); CK(dalloc(t.p, 3 * (size_t)cap)); CK(dalloc(t.mu, cap));
    t.cap = cap;
  }
  t.nc = nc;
  t.colour_off = ctx->colour_off;
  CK(hipMemcpyAsync(t.pos, pos_before, 3 * D * n, hipMemcpyDeviceToDevice, s));
  CK(hipMemcpyAsync(t.sb, ctx->sb, (size_t)kSB * D * n, hipMemcpyDeviceToDevice, s));
  adj_pack_vel((uint32_t)n, ctx->v, ctx->w, t.vel, s);
  if (nc) {
    CK(hipMemcpyAsync(t.b1, ctx->sol.b1, 4 * (size_t)nc, hipMemcpyDeviceToDevice, s));
    CK(hipMemcpyAsync(t.b2, ctx->sol.b2, 4 * (size_t)nc, hipMemcpyDeviceToDevice, s));
    adj_pack_p(nc, ctx->sol.f, ctx->sol.cap, t.p, t.mu, s);
  }
  if (nc > a->ccap) {
    dfree(a->jac); dfree(a->minv); dfree(a->lambda);
    a->ccap = 0;
    const uint32_t cap = std::max<uint32_t>(nc + nc / 4, 1024);
    CK(dalloc(a->jac, (size_t)cap * ad::kAdjIn * ad::kAdjOut));
    CK(dalloc(a->minv, 9 * (size_t)cap));
    CK(dalloc(a->lambda, 3 * (size_t)cap));
    a->ccap = cap;
  }
  return PEHIP_OK;
}

// Reverse of one recorded step: (xbar, vbar, wbar) after the step -> before it; mubar accumulates.
static pehip_status reverse_step(pehip_ctx* ctx, AdjointState* a, const StepTape& t, uint32_t sweeps, double tol,
                                 uint32_t* used, double* last_delta) {
  const hipStream_t s = ctx->stream;
  const uint32_t n = ctx->n;
  adj_init(n, ctx->fixed, a->xbar, a->vbar, a->wbar, a->dt, a->alpha, s);
  AdjTape T{t.nc, t.b1, t.b2, t.p, t.mu, t.pos, t.sb, t.vel};
  AdjConsts A{ctx->shape, ctx->param, ctx->pnormal, ctx->pd, ctx->pos, a->prm.erp, 1.0 / a->dt};
  *used = 0;
  *last_delta = 0;
  if (t.nc) {
    fill_u64((unsigned long long*)a->lambda, 0ull, 3 * (size_t)t.nc, s);
    adj_jacobian(a->prm.relaxation_model, T, A, a->jac, s);
    adj_prepare(t.nc, a->jac, a->minv, s);
    const uint32_t ncol = (uint32_t)t.colour_off.size() - 1;
    for (uint32_t it = 0; it < sweeps; ++it) {
      fill_u64(a->delta, 0ull, 1, s);
      for (uint32_t k = 0; k < ncol; ++k)
        adj_sweep(T, t.colour_off[k], t.colour_off[k + 1], a->jac, a->minv, a->alpha, a->lambda, a->delta, s);
      *used = it + 1;
      if (tol > 0 && (it % 8 == 7 || it + 1 == sweeps)) {
        CK(hipMemcpyAsync(ctx->hstage, a->delta, 8, hipMemcpyDeviceToHost, s));
        CKS(sync_check(ctx));
        std::memcpy(last_delta, ctx->hstage, 8);
        if (*last_delta < tol) break;
      }
    }
    if (tol <= 0) {
      CK(hipMemcpyAsync(ctx->hstage, a->delta, 8, hipMemcpyDeviceToHost, s));
      CKS(sync_check(ctx));
      std::memcpy(last_delta, ctx->hstage, 8);
    }
    adj_assemble(T, a->jac, a->alpha, a->lambda, a->xbar, a->mubar, s);
  }
  adj_finish(n, a->alpha, a->vbar, a->wbar, s);
  return PEHIP_OK;
}

extern "C" {

pehip_status pehip_adjoint_forward(pehip_ctx* ctx, double dt, uint32_t nsteps, uint32_t checkpoint_every) {
  if (!ctx) return PEHIP_INVALID_ARG;
  if (!(dt > 0)) return fail(ctx, PEHIP_INVALID_ARG, "Invalid time step size");
  if (ctx->n == 0) return fail(ctx, PEHIP_INVALID_ARG, "no bodies uploaded");
  if (nsteps == 0 || checkpoint_every == 0) return fail(ctx, PEHIP_INVALID_ARG, "nsteps and checkpoint_every must be > 0");
  CKS(check_supported(ctx, ctx->prm));
  hipSetDevice(ctx->device);
  CK(hipStreamSynchronize(ctx->stream));
  pe_adjoint_free(ctx);
  AdjointState* a = new AdjointState();
  ctx->adj = a;
  a->dt = dt;
  a->nsteps = nsteps;
  a->every = checkpoint_every;
  a->prm = ctx->prm;
  const size_t n = ctx->n;
  const pehip_status result = [&]() -> pehip_status {
    for (uint32_t st = 0; st < nsteps; ++st) {
      if (st % checkpoint_every == 0) {
        double* c = nullptr;
        CK(dalloc(c, kStateDoubles * n));
        a->checkpoints.push_back(c);
        save_state(ctx, c);
      }
      CKS(pe_one_step(ctx, dt));
    }
    CK(dalloc(a->final_state, kStateDoubles * n));
    save_state(ctx, a->final_state);
    return sync_check(ctx);
  }();
  // a run that did not complete leaves nothing to differentiate
  if (result != PEHIP_OK) pe_adjoint_free(ctx);
  return result;
}

pehip_status pehip_adjoint_backward(pehip_ctx* ctx, const double* dL_dpos, const double* dL_dv, const double* dL_dw,
                                    double* dL_dpos0, double* dL_dv0, double* dL_dw0, double* dL_dmu,
                                    uint32_t adjoint_sweeps, double adjoint_tol, pehip_adjoint_stats* stats) {
  if (!ctx) return PEHIP_INVALID_ARG;
  AdjointState* a = ctx->adj;
  if (!a) return fail(ctx, PEHIP_INVALID_ARG, "no differentiable run: call pehip_adjoint_forward first");
  if (adjoint_sweeps == 0) return fail(ctx, PEHIP_INVALID_ARG, "adjoint_sweeps must be > 0");
  hipSetDevice(ctx->device);
  const hipStream_t s = ctx->stream;
  const size_t n = ctx->n, D = sizeof(double);
  if (!a->xbar) CK(dalloc(a->xbar, 3 * n));
  if (!a->vbar) CK(dalloc(a->vbar, 3 * n));
  if (!a->wbar) CK(dalloc(a->wbar, 3 * n));
  if (!a->alpha) CK(dalloc(a->alpha, 6 * n));
  if (!a->mubar) CK(dalloc(a->mubar, 1));
  if (!a->delta) CK(dalloc(a->delta, 1));
  fill_u64((unsigned long long*)a->xbar, 0ull, 3 * n, s);
  fill_u64((unsigned long long*)a->vbar, 0ull, 3 * n, s);
  fill_u64((unsigned long long*)a->wbar, 0ull, 3 * n, s);
  fill_u64((unsigned long long*)a->mubar, 0ull, 1, s);
  CK(hipStreamSynchronize(s));
  if (dL_dpos) CK(hipMemcpy(a->xbar, dL_dpos, 3 * D * n, hipMemcpyHostToDevice));
  if (dL_dv) CK(hipMemcpy(a->vbar, dL_dv, 3 * D * n, hipMemcpyHostToDevice));
  if (dL_dw) CK(hipMemcpy(a->wbar, dL_dw, 3 * D * n, hipMemcpyHostToDevice));
  // the re-runs use the forward run's parameters; the caller's come back however the pass ends
  const pehip_params saved = ctx->prm;
  ctx->prm = a->prm;
  const bool prof = ctx->profiling;
  ctx->profiling = false;
  pehip_adjoint_stats st{};
  double* pos_before = nullptr;
  pehip_status result = [&]() -> pehip_status {
    CK(dalloc(pos_before, 3 * n));
    const uint32_t nseg = (uint32_t)a->checkpoints.size();
    if (a->tape.size() < a->every) a->tape.resize(a->every);
    for (int seg = (int)nseg - 1; seg >= 0; --seg) {
      const uint32_t s0 = (uint32_t)seg * a->every;
      const uint32_t m = std::min(a->every, a->nsteps - s0);
      restore_state(ctx, a->checkpoints[seg]);
      for (uint32_t j = 0; j < m; ++j) {
        CK(hipMemcpyAsync(pos_before, ctx->pos, 3 * D * n, hipMemcpyDeviceToDevice, s));
        CKS(pe_one_step(ctx, a->dt));
        CKS(record(ctx, a, a->tape[j], pos_before));
        st.max_contacts = std::max<uint64_t>(st.max_contacts, ctx->nc);
      }
      for (int j = (int)m - 1; j >= 0; --j) {
        uint32_t used = 0;
        double d = 0;
        CKS(reverse_step(ctx, a, a->tape[j], adjoint_sweeps, adjoint_tol, &used, &d));
        st.total_sweeps += used;
        st.max_last_delta = std::max(st.max_last_delta, d);
        ++st.steps;
      }
    }
    return PEHIP_OK;
  }();
  // the world is left in the forward run's final state whether or not the pass succeeded
  restore_state(ctx, a->final_state);
  const hipError_t se = hipStreamSynchronize(s);
  dfree(pos_before);
  if (result == PEHIP_OK) {
    result = [&]() -> pehip_status {
      CK(se);
      if (dL_dpos0) CK(hipMemcpy(dL_dpos0, a->xbar, 3 * D * n, hipMemcpyDeviceToHost));
      if (dL_dv0) CK(hipMemcpy(dL_dv0, a->vbar, 3 * D * n, hipMemcpyDeviceToHost));
      if (dL_dw0) CK(hipMemcpy(dL_dw0, a->wbar, 3 * D * n, hipMemcpyDeviceToHost));
      if (dL_dmu) CK(hipMemcpy(dL_dmu, a->mubar, D, hipMemcpyDeviceToHost));
      if (stats) *stats = st;
      return sync_check(ctx);
    }();
  }
  ctx->prm = saved;
  ctx->profiling = prof;
  return result;
}

}  // extern "C"
