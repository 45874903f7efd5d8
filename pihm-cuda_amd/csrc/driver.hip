// driver.hip - MM-PIHM model-step loop and the device backend.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/pihm_driver.h"
#include "../../include/pihm_bc.h"
#include "../../include/pihm_lsm.h"
#include "../../include/pihm_output.h"
#include "../../include/pihm_noahsoil.h"
#include "../../include/pihm_noahlsm.h"
#include "common.hpp"
#include "ctx.hpp"
#include "nvector_impl.hpp"

using namespace pihm;

static int gpu_rhs(realtype t, N_Vector y, N_Vector ydot, void *ud);
static PihmRhsFn gpu_rhs_fn() { return gpu_rhs; }

extern "C" int pihm_drive(const pihm_backend *be, N_Vector y, const pihm_run_config *cfg,
                          pihm_run_stats *st) {
  if (!be || !y || !cfg || !be->rhs) return PIHM_ERR_ARG;
  if (!(cfg->stepsize > 0.0) || !(cfg->etstep > 0.0) || cfg->t_end < cfg->t_start) return PIHM_ERR_ARG;
  const auto t0w = std::chrono::steady_clock::now();
  pihm_run_stats local;
  std::memset(&local, 0, sizeof(local));

  // SetCVodeParam: src/ode.c:413-529
  void *cv = PihmCVodeCreate(PIHM_CV_BDF, PIHM_CV_NEWTON);
  if (!cv) return PIHM_ERR_ARG;
  double maxstep = cfg->stepsize;
  int flag = PihmCVodeInit(cv, be->rhs, 0.0, y);
  if (!flag) flag = PihmCVodeSStolerances(cv, cfg->reltol, cfg->abstol);
  if (!flag) flag = PihmCVodeSetUserData(cv, be->rhs_data);
  if (!flag) flag = PihmCVodeSetInitStep(cv, cfg->initstep);
  if (!flag) flag = PihmCVodeSetStabLimDet(cv, 1);
  if (!flag) flag = PihmCVodeSetMaxStep(cv, maxstep);
  if (!flag) flag = PihmCVodeSetMaxNumSteps(cv, cfg->max_steps_per_call > 0 ? cfg->max_steps_per_call
                                                                            : (long)(cfg->stepsize * 10));
  const bool pre = cfg->precond && be->prec_setup && be->prec_solve;
  if (cfg->precond && !pre) flag = PIHM_ERR_ARG;
  if (!flag) flag = PihmCVSpgmr(cv, pre ? PIHM_PREC_RIGHT : PIHM_PREC_NONE, cfg->maxl);
  if (!flag && pre) flag = PihmCVSpilsSetPreconditioner(cv, be->prec_setup, be->prec_solve);
  if (!flag && cfg->rec_cap > 0)
    flag = PihmCVodeSetStepRecord(cv, cfg->rec_h, cfg->rec_q, cfg->rec_iters, cfg->rec_cap, cfg->rec_n);
  if (!flag && cfg->replay_n > 0)
    flag = PihmCVodeSetStepReplay(cv, cfg->replay_h, cfg->replay_q, cfg->replay_iters, cfg->replay_n);
  if (!flag && cfg->graphs && be->rhs == gpu_rhs_fn()) flag = PihmCVodeSetGraphRhs(cv, 1);
  if (flag) { PihmCVodeFree(&cv); return flag; }

  long nst0 = 0, ncfn0 = 0, nni0 = 0;
  const double tstart = cfg->t_start;
  const long nsteps = (long)std::llround((cfg->t_end - cfg->t_start) / cfg->stepsize);
  const long et_every = std::max(1L, (long)std::llround(cfg->etstep / cfg->stepsize));
  for (long k = 0; k < nsteps; k++) {
    const double t = tstart + k * cfg->stepsize;
    // ApplyBc every model step (src/pihm.c:17-24)
    if (be->apply_bc && (flag = be->apply_bc(be->user, t))) break;
    // ApplyForc every LSM step (src/pihm.c:30-45)
    if (k % et_every == 0 && be->apply_forcing) {
      if ((flag = be->apply_forcing(be->user, t))) break;
    }
    // SolveCVode: src/ode.c:555-599 (solver time runs from 0 at t_start)
    const double tout = (double)((k + 1) * cfg->stepsize);
    double solvert = 0.0;
    flag = PihmCVodeSetStopTime(cv, tout);
    if (flag) break;
    flag = PihmCVode(cv, tout, y, &solvert, PIHM_CV_NORMAL);
    if (flag < 0) break;
    // Summary: ws0 <- ws (src/update.c:46)
    if (be->summary && (flag = be->summary(be->user, y))) break;
    local.model_steps++;
    if (be->model_step_end && (flag = be->model_step_end(be->user, tstart + tout, y))) break;
    // AdjCVodeMaxStep: src/ode.c:601-661
    long nst, ncfn, nni;
    PihmCVodeGetNumSteps(cv, &nst);
    PihmCVodeGetNumNonlinSolvConvFails(cv, &ncfn);
    PihmCVodeGetNumNonlinSolvIters(cv, &nni);
    const double dsteps = (double)(nst - nst0);
    const double nfails = (double)(ncfn - ncfn0) / dsteps;
    const double niters = (double)(nni - nni0) / dsteps;
    if (nfails > cfg->nncfn || niters >= cfg->nnimax) maxstep /= cfg->decr;
    if (nfails == 0.0 && niters <= cfg->nnimin) maxstep *= cfg->incr;
    maxstep = (maxstep < cfg->stepsize) ? maxstep : cfg->stepsize;
    maxstep = (maxstep > cfg->stmin) ? maxstep : cfg->stmin;
    PihmCVodeSetMaxStep(cv, maxstep);
    nst0 = nst; ncfn0 = ncfn; nni0 = nni;
  }
  if (flag > 0) flag = 0;   // CV_TSTOP_RETURN is the normal outcome of each step
  PihmCVodeGetNumSteps(cv, &local.nst);
  PihmCVodeGetNumRhsEvals(cv, &local.nfe);
  PihmCVodeGetNumJtimesRhsEvals(cv, &local.nfes);
  PihmCVodeGetNumNonlinSolvIters(cv, &local.nni);
  PihmCVodeGetNumLinIters(cv, &local.nli);
  PihmCVodeGetNumNonlinSolvConvFails(cv, &local.ncfn);
  PihmCVodeGetNumErrTestFails(cv, &local.netf);
  PihmCVodeGetNumLinSolvSetups(cv, &local.nsetups);
  PihmCVSpilsGetNumPrecEvals(cv, &local.npe);
  PihmCVSpilsGetNumPrecSolves(cv, &local.nps);
  PihmCVodeFree(&cv);
  local.flag = flag;
  local.wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0w).count();
  if (st) *st = local;
  return flag;
}

// ---------------------------------------------------------------- device backend
namespace {

__global__ void k_forcing(int ne, const int32_t *station, const double *prcp, const double *edir,
                          const double *ett, int nrec, double rec_dt, double t, double *pcpdrp,
                          double *oedir, double *oett) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ne) return;
  long k = (long)floor(t / rec_dt);
  if (k < 0) k = 0;
  if (k > nrec - 2) k = nrec - 2;
  const double ta = (double)k * rec_dt, tb = (double)(k + 1) * rec_dt;
  const long s = station[p];
  const double *P = prcp + s * (long)nrec;
  // IntrplForc: src/forcing.c:514-519
  pcpdrp[p] = ((tb - t) * P[k] + (t - ta) * P[k + 1]) / (tb - ta);
  oedir[p] = ((tb - t) * edir[k] + (t - ta) * edir[k + 1]) / (tb - ta);
  oett[p] = ((tb - t) * ett[k] + (t - ta) * ett[k + 1]) / (tb - ta);
}

__global__ void k_copy(long n, const double *src, double *dst) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace

struct pihm_gpu_backend {
  pihm_ctx *ctx;
  int nrec;
  double rec_dt;
  int32_t *station;
  double *prcp, *edir, *ett;
  pihm_noah_soil *noah_soil;     // the _NOAH_ build's soil-water step at each Summary
  int summaries;
};

static int gpu_rhs(realtype t, N_Vector y, N_Vector ydot, void *ud) {
  return PIHM_ODE_Hip(t, y, ydot, ud);
}
// gpu_rhs only enqueues work on the context's stream once the NaN check is off
// (pihm_run_gpu), so the integrator may replay it from a graph.

static int gpu_forcing(void *user, double t) {
  pihm_gpu_backend *b = (pihm_gpu_backend *)user;
  pihm_ctx *c = b->ctx;
  hipLaunchKernelGGL(k_forcing, dim3(div_up(c->ne, kBlock)), dim3(kBlock), 0, c->stream, c->ne,
                     b->station, b->prcp, b->edir, b->ett, b->nrec, b->rec_dt, t,
                     c->args.e.pcpdrp, c->args.e.edir, c->args.e.ett);
  if (hipGetLastError() != hipSuccess) return PIHM_ERR_HIP;
  return pihm_ctx_noah_refresh(c);
}

// ws0.surf <- y[SURF] (device order on both sides)
static int copy_ws0(pihm_ctx *c, N_Vector y) {
  hipLaunchKernelGGL(k_copy, dim3(div_up(c->ne, kBlock)), dim3(kBlock), 0, c->stream, (long)c->ne,
                     N_VGetDeviceArrayPointer_Hip(y), c->args.e.ws0s);
  return hipGetLastError() == hipSuccess ? 0 : PIHM_ERR_HIP;
}

static int gpu_summary(void *user, N_Vector y) {
  pihm_gpu_backend *b = (pihm_gpu_backend *)user;
  if (b->noah_soil) {
    // Summary's MassBalance and NoahHydrol (src/pihm.c:95-110); the first call
    // is the initial state, whose storage starts the balance
    const double *yd = N_VGetDeviceArrayPointer_Hip(y);
    const int rc = (b->summaries++ == 0) ? pihm_noah_soil_start(b->noah_soil, yd)
                                         : pihm_noah_soil_summary(b->noah_soil, b->ctx->dt, yd);
    if (rc) return rc;
  }
  return copy_ws0(b->ctx, y);
}

extern "C" int pihm_gpu_backend_set_noah_soil(pihm_gpu_backend *b, pihm_noah_soil *s) {
  if (!b) return PIHM_ERR_ARG;
  b->noah_soil = s;
  b->summaries = 0;
  return PIHM_OK;
}

extern "C" int pihm_gpu_backend_create(pihm_ctx *c, const pihm_forcing *f, pihm_gpu_backend **out) {
  if (!c || !f || !out || f->nrec < 2 || f->nstation < 1 || !(f->rec_dt > 0.0)) return PIHM_ERR_ARG;
  if (!f->prcp || !f->edir || !f->ett || !f->station) return PIHM_ERR_ARG;
  for (int i = 0; i < c->ne; i++)
    if (f->station[i] < 0 || f->station[i] >= f->nstation) return PIHM_ERR_ARG;
  pihm_gpu_backend *b = new pihm_gpu_backend();
  b->ctx = c;
  b->nrec = f->nrec;
  b->rec_dt = f->rec_dt;
  std::vector<int32_t> st(c->ne);
  for (int p = 0; p < c->ne; p++) st[p] = f->station[c->perm[p]];
  PIHM_HIP_TRY(hipMalloc(&b->station, c->ne * sizeof(int32_t)));
  PIHM_HIP_TRY(hipMalloc(&b->prcp, (size_t)f->nstation * f->nrec * sizeof(double)));
  PIHM_HIP_TRY(hipMalloc(&b->edir, f->nrec * sizeof(double)));
  PIHM_HIP_TRY(hipMalloc(&b->ett, f->nrec * sizeof(double)));
  PIHM_HIP_TRY(hipMemcpy(b->station, st.data(), c->ne * sizeof(int32_t), hipMemcpyHostToDevice));
  PIHM_HIP_TRY(hipMemcpy(b->prcp, f->prcp, (size_t)f->nstation * f->nrec * sizeof(double), hipMemcpyHostToDevice));
  PIHM_HIP_TRY(hipMemcpy(b->edir, f->edir, f->nrec * sizeof(double), hipMemcpyHostToDevice));
  PIHM_HIP_TRY(hipMemcpy(b->ett, f->ett, f->nrec * sizeof(double), hipMemcpyHostToDevice));
  *out = b;
  return PIHM_OK;
}

extern "C" void pihm_gpu_backend_destroy(pihm_gpu_backend *b) {
  if (!b) return;
  (void)hipFree(b->station);
  (void)hipFree(b->prcp);
  (void)hipFree(b->edir);
  (void)hipFree(b->ett);
  delete b;
}

extern "C" int pihm_gpu_backend_get(pihm_gpu_backend *b, pihm_backend *out) {
  if (!b || !out) return PIHM_ERR_ARG;
  std::memset(out, 0, sizeof(*out));
  out->rhs = gpu_rhs;
  out->rhs_data = b->ctx;
  out->apply_forcing = gpu_forcing;
  out->summary = gpu_summary;
  out->user = b;
  return PIHM_OK;
}

// Device vector, stream and NaN polling around pihm_drive for a backend whose
// rhs_data is the context.  The backend's summary sees the device vector first
// (initial ws0 = initial state, src/initialize.c InitVar).
static int run_device(pihm_ctx *c, const pihm_run_config *cfg, double *y_inout, pihm_run_stats *st,
                      const pihm_backend *be) {
  // graphs need a stream of their own (capture is not possible on the null stream)
  hipStream_t own = nullptr;
  void *saved_stream = c->stream;
  if (cfg->graphs && !c->stream) {
    if (hipStreamCreate(&own) != hipSuccess) return PIHM_ERR_HIP;
    c->stream = own;
  }
  N_Vector y = N_VNew_Hip(c->neq);
  if (!y) {
    c->stream = (hipStream_t)saved_stream;
    if (own) (void)hipStreamDestroy(own);
    return PIHM_ERR_HIP;
  }
  N_VSetStream_Hip(y, c->stream);
  const int saved_check = c->nan_check;
  c->nan_check = 0;   // NaN shows up in the norms; polled once at the end
  int rc = pihm_ctx_state_to_device(c, y_inout, y);
  // a flag an earlier call left on this context is not this run's
  if (!rc && (rc = pihm_ctx_poll_nan(c)) == PIHM_ERR_NAN) rc = PIHM_OK;
  if (!rc) rc = be->summary(be->user, y);
  if (!rc) rc = pihm_drive(be, y, cfg, st);
  {
    // Read and clear the flag whatever the run returned: a NaN that made the
    // integrator give up is reported as the NaN, and the flag is not left set
    // for the next call on the context.
    const int rn = pihm_ctx_poll_nan(c);
    if (rn == PIHM_ERR_NAN || (!rc && rn)) rc = rn;
  }
  if (!rc || rc == PIHM_ERR_NAN) {
    int rc2 = pihm_ctx_state_from_device(c, y, y_inout);
    if (!rc) rc = rc2;
  }
  c->nan_check = saved_check;
  N_VDestroy_Hip(y);
  if (own) {
    (void)hipStreamSynchronize(own);
    c->stream = (hipStream_t)saved_stream;
    (void)hipStreamDestroy(own);
  }
  return rc;
}

extern "C" int pihm_run_gpu(pihm_ctx *c, const pihm_forcing *f, const pihm_run_config *cfg,
                            double *y_inout, pihm_run_stats *st) {
  if (!c || !f || !cfg || !y_inout) return PIHM_ERR_ARG;
  pihm_gpu_backend *b = nullptr;
  int rc = pihm_gpu_backend_create(c, f, &b);
  if (rc) return rc;
  pihm_backend be;
  pihm_gpu_backend_get(b, &be);
  rc = run_device(c, cfg, y_inout, st, &be);
  pihm_gpu_backend_destroy(b);
  return rc;
}

extern "C" int pihm_run_gpu_noah(pihm_ctx *c, const pihm_forcing *f, pihm_noah_soil *soil,
                                 const pihm_run_config *cfg, double *y_inout, pihm_run_stats *st) {
  if (!c || !f || !soil || !cfg || !y_inout || !c->noah) return PIHM_ERR_ARG;
  pihm_gpu_backend *b = nullptr;
  int rc = pihm_gpu_backend_create(c, f, &b);
  if (rc) return rc;
  pihm_gpu_backend_set_noah_soil(b, soil);
  pihm_backend be;
  pihm_gpu_backend_get(b, &be);
  rc = run_device(c, cfg, y_inout, st, &be);
  pihm_gpu_backend_destroy(b);
  return rc;
}

// ---------------------------------------------------------------- producer-driven runs
namespace pihm {
pihm_ctx *lsm_context(const pihm_lsm *l);   // lsm.hip
pihm_ctx *bc_context(const pihm_bc *b);     // bc.hip
}

namespace {

// Every LSM step the land-surface step runs on the state the last Summary saw
// (IntcpSnowEt reads elem.ws, src/pihm.c:30-63); the driver integrates y in
// place, so the vector Summary is handed is the one the next step reads.
struct ProducersRun {
  pihm_ctx *ctx;
  pihm_producers p;
  N_Vector y;
  double etstep, t_start;
};

int run_bc(void *user, double t) {
  ProducersRun *r = (ProducersRun *)user;
  return pihm_bc_apply(r->p.bc, t);
}

int run_lsm(void *user, double t) {
  ProducersRun *r = (ProducersRun *)user;
  int rc = pihm_lsm_step(r->p.lsm, t, r->etstep, r->y);
  if (!rc && r->p.out) rc = pihm_output_update(r->p.out, PIHM_OUT_LS_STEP, r->y);   // src/pihm.c:70-72
  return rc;
}

int run_summary(void *user, N_Vector y) {
  ProducersRun *r = (ProducersRun *)user;
  if (!r->y && r->p.out) {
    // the first call is the initial state (ws0 = ws after initialisation)
    const int rc = pihm_output_start(r->p.out, y);
    if (rc) return rc;
  }
  r->y = y;
  return copy_ws0(r->ctx, y);
}

// UpdPrintVar(HYDROL_STEP) and PrintData after Summary (src/pihm.c:112-160)
int run_step_end(void *user, double t, N_Vector y) {
  ProducersRun *r = (ProducersRun *)user;
  int rc = pihm_output_update(r->p.out, PIHM_OUT_HYDROL_STEP, y);
  if (!rc) rc = pihm_output_print(r->p.out, t, t - r->t_start);
  return rc;
}

}  // namespace

extern "C" int pihm_run_gpu_producers(pihm_ctx *c, const pihm_producers *p, const pihm_run_config *cfg,
                                      double *y_inout, pihm_run_stats *st) {
  if (!c || !p || !cfg || !y_inout) return PIHM_ERR_ARG;
  if ((p->lsm && pihm::lsm_context(p->lsm) != c) || (p->bc && pihm::bc_context(p->bc) != c)) return PIHM_ERR_ARG;
  ProducersRun r{c, *p, nullptr, cfg->etstep, cfg->t_start};
  pihm_backend be;
  std::memset(&be, 0, sizeof(be));
  be.rhs = gpu_rhs;
  be.rhs_data = c;
  be.apply_forcing = p->lsm ? run_lsm : nullptr;
  be.summary = run_summary;
  be.apply_bc = p->bc ? run_bc : nullptr;
  be.model_step_end = p->out ? run_step_end : nullptr;
  be.user = &r;
  return run_device(c, cfg, y_inout, st, &be);
}

extern "C" int pihm_run_gpu_lsm(pihm_ctx *c, pihm_lsm *lsm, pihm_bc *bc, const pihm_run_config *cfg,
                                double *y_inout, pihm_run_stats *st) {
  if (!lsm) return PIHM_ERR_ARG;
  const pihm_producers p{lsm, bc, nullptr};
  return pihm_run_gpu_producers(c, &p, cfg, y_inout, st);
}

// ---------------------------------------------------------------- Flux-PIHM runs
namespace pihm {
pihm_noah_soil *noah_lsm_soil(pihm_noah_lsm *n);   // noahlsm.hip
pihm_ctx *noah_lsm_context(pihm_noah_lsm *n);
}

namespace {

// The _NOAH_ model-step loop (src/pihm.c:10-187): Noah every LSM step on the
// soil column the last soil-water step left, Summary with MassBalance and the
// soil-water step after every model step.
struct FluxRun {
  pihm_ctx *ctx;
  pihm_noah_lsm *lsm;
  pihm_noah_soil *soil;
  double etstep;
  int summaries;
};

int flux_forcing(void *user, double t) {
  FluxRun *r = (FluxRun *)user;
  return pihm_noah_lsm_step(r->lsm, t, r->etstep);
}

int flux_summary(void *user, N_Vector y) {
  FluxRun *r = (FluxRun *)user;
  const double *yd = N_VGetDeviceArrayPointer_Hip(y);
  const int rc = (r->summaries++ == 0) ? pihm_noah_soil_start(r->soil, yd)
                                       : pihm_noah_soil_summary(r->soil, r->ctx->dt, yd);
  if (rc) return rc;
  return copy_ws0(r->ctx, y);
}

}  // namespace

extern "C" int pihm_run_gpu_flux(pihm_noah_lsm *n, const pihm_run_config *cfg, double *y_inout,
                                 pihm_run_stats *st) {
  if (!n || !cfg || !y_inout) return PIHM_ERR_ARG;
  pihm_ctx *c = pihm::noah_lsm_context(n);
  if (!c || !c->noah) return PIHM_ERR_ARG;
  FluxRun r{c, n, pihm::noah_lsm_soil(n), cfg->etstep, 0};
  pihm_backend be;
  std::memset(&be, 0, sizeof(be));
  be.rhs = gpu_rhs;
  be.rhs_data = c;
  be.apply_forcing = flux_forcing;
  be.summary = flux_summary;
  be.user = &r;
  return run_device(c, cfg, y_inout, st, &be);
}
