/*
 * pihm_driver.h - model-step loop of MM-PIHM around the integrator.
 *
 * pihm_drive restates the reference's time loop (src/main.c:217-254) and one
 * model step (src/pihm.c:10-187): apply forcing every LSM step, integrate to
 * the next model step with the stop time set (SolveCVode, src/ode.c:555-599),
 * copy the state into ws0 (Summary, src/update.c:3-99), then adapt the
 * maximum step (AdjCVodeMaxStep, src/ode.c:601-661).  The right-hand side,
 * forcing and summary are supplied by a backend, so the same loop drives the
 * device path and a host implementation.
 *
 * Forcing: per-station precipitation and domain-wide soil evaporation and
 * transpiration records at a fixed spacing, interpolated linearly in time
 * like IntrplForc (src/forcing.c:482-530); pcpdrp is the interpolated
 * precipitation (interception is not modelled), edir/ett the interpolated
 * evaporation and transpiration.
 */
#ifndef PIHM_DRIVER_H
#define PIHM_DRIVER_H

#include <stdint.h>
#include "pihm_nvector.h"
#include "pihm_cvode.h"
#include "pihm_rhs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pihm_run_config {
  double t_start, t_end;        /* seconds */
  double stepsize;              /* MODEL_STEPSIZE */
  double etstep;                /* LSM_STEP, a multiple of stepsize */
  double reltol, abstol, initstep;
  double nncfn, nnimax, nnimin, decr, incr, stmin;   /* NUM_NONCOV_FAIL .. MIN_MAXSTEP */
  int maxl;                     /* SPGMR dimension (0 = 5) */
  long max_steps_per_call;      /* CVodeSetMaxNumSteps; 0 = stepsize*10 as the reference */
  int graphs;                   /* pihm_run_gpu: replay Krylov iterations from HIP graphs */
  int precond;                  /* 0: SPGMR without preconditioning, as the reference
                                   (src/ode.c:523); 1: right block-Jacobi preconditioner
                                   from the backend's prec_setup / prec_solve */
  /* optional step record / replay (PihmCVodeSetStepRecord / SetStepReplay) */
  double *rec_h;
  int32_t *rec_q, *rec_iters;   /* rec_iters may be NULL */
  long rec_cap, *rec_n;
  const double *replay_h;
  const int32_t *replay_q, *replay_iters;   /* replay_iters may be NULL */
  long replay_n;
} pihm_run_config;

typedef struct pihm_forcing {
  int32_t nstation, nrec;
  double rec_dt;                /* record spacing [s], records at k*rec_dt */
  const double *prcp;           /* [nstation*nrec] m/s */
  const double *edir, *ett;     /* [nrec] m/s */
  const int32_t *station;       /* [nelem], caller order */
} pihm_forcing;

typedef struct pihm_run_stats {
  long nst, nfe, nfes, nni, nli, ncfn, netf;
  long model_steps;
  double wall_s;
  int flag;                     /* last integrator flag */
  long nsetups, npe, nps;       /* preconditioner setups, Jacobian evaluations, solves */
} pihm_run_stats;

/* Backend: the right-hand side plus the per-step producers. */
typedef struct pihm_backend {
  PihmRhsFn rhs;
  void *rhs_data;
  int (*apply_forcing)(void *user, double t);   /* sets pcpdrp/edir/ett for time t */
  int (*summary)(void *user, N_Vector y);       /* ws0 <- state after a model step */
  int (*model_step_end)(void *user, double t, N_Vector y);  /* optional, may be NULL */
  void *user;
  int (*apply_bc)(void *user, double t);        /* every model step before the forcing
                                                   (ApplyBc, src/pihm.c:17-24); may be NULL */
  PihmPrecSetupFn prec_setup;                   /* used when cfg->precond; called with rhs_data */
  PihmPrecSolveFn prec_solve;
} pihm_backend;

int pihm_drive(const pihm_backend *be, N_Vector y, const pihm_run_config *cfg,
               pihm_run_stats *stats);

/* Device backend for a context: forcing records live in HBM and pcpdrp,
 * edir, ett and ws0 are produced on the device (no per-step uploads). */
typedef struct pihm_gpu_backend pihm_gpu_backend;
int  pihm_gpu_backend_create(pihm_ctx *ctx, const pihm_forcing *forc, pihm_gpu_backend **out);
void pihm_gpu_backend_destroy(pihm_gpu_backend *b);
int  pihm_gpu_backend_get(pihm_gpu_backend *b, pihm_backend *out);

/* One call: device vector, backend and driver for a whole run.
 * y_inout is the state in caller order. */
int pihm_run_gpu(pihm_ctx *ctx, const pihm_forcing *forc, const pihm_run_config *cfg,
                 double *y_inout, pihm_run_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
