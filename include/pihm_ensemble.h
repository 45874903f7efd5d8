/*
 * pihm_ensemble.h - many MM-PIHM solves advanced together on one device.
 *
 * An ensemble runs every member of a batch context (pihm_ctx_create_batch)
 * through the model-step loop of pihm_drive (src/main.c:217-254,
 * src/pihm.c:10-187): forcing every LSM step, integration to the next model
 * step with the stop time set, ws0 <- state, and AdjCVodeMaxStep
 * (src/ode.c:601-661).  Each member has its own integrator state (order, step
 * size, Nordsieck history, counters) and takes its own branches; the
 * ensemble interleaves the members so that the vector operations and
 * right-hand-side evaluations of all members that are ready run as one launch
 * per kind, with one host synchronisation per round of inner products.
 *
 * The integrator per member is the BDF/Newton/SPGMR method of pihm_cvode.h
 * (same steps, same decisions).  Its inner products are summed per member in
 * a fixed order that depends only on that member's length, so a member's
 * trajectory is bit-identical whichever other members run beside it.  The
 * ensemble requires an initial step (cfg->initstep > 0), as MM-PIHM always
 * sets one (src/ode.c:480).
 *
 * Failure is per member: a NaN in a member's derivative, or an integrator
 * error, retires that member (its flag in pihm_ens_member_stats) and the
 * other members continue.
 *
 * The reference has no ensemble; the entry points here drive many copies of
 * the path the reference runs once (CVodeInit/SetCVodeParam, src/ode.c:413-529,
 * then the time loop, src/main.c:217-254).
 */
#ifndef PIHM_ENSEMBLE_H
#define PIHM_ENSEMBLE_H

#include "pihm_driver.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pihm_ensemble pihm_ensemble;

/* forc: one forcing description per member (station indices in the member's
 * caller order); cfg: shared configuration (t_start and t_end are unused;
 * the solver time of every member starts at 0).  The ensemble allocates its
 * vectors (20 state-length vectors per member) in HBM. */
int  pihm_ens_create(pihm_ctx *batch, const pihm_forcing *forc, const pihm_run_config *cfg,
                     pihm_ensemble **out);
void pihm_ens_destroy(pihm_ensemble *e);

/* Initial state of member k in its caller order; resets the member's
 * integrator, its model-step count, its carried river-edge fluxes, and sets
 * ws0 = the surface state (InitVar). */
int  pihm_ens_set_state(pihm_ensemble *e, int k, const double *y);
/* State of member k after its last model step, caller order. */
int  pihm_ens_get_state(pihm_ensemble *e, int k, double *y);
/* Members excluded from runs keep their state. */
int  pihm_ens_set_member_enabled(pihm_ensemble *e, int k, int on);

/* Advance every enabled member that has not failed to solver time t_end
 * (a multiple of cfg->stepsize).  Returns PIHM_OK when every such member got
 * there, else the number of members that failed during this call. */
int  pihm_ens_run(pihm_ensemble *e, double t_end);

/* Counters of member k (flag: 0, or the error that retired the member:
 * an integrator flag < 0, or PIHM_ERR_NAN); model_steps: model steps done;
 * wall_s: unused (0). */
int  pihm_ens_member_stats(pihm_ensemble *e, int k, pihm_run_stats *st);

/* Host synchronisations ("ticks"), vector-operation launches and batched
 * right-hand-side evaluations since creation. */
int  pihm_ens_counters(pihm_ensemble *e, long *ticks, long *op_launches, long *rhs_launches);

#ifdef __cplusplus
}
#endif
#endif
