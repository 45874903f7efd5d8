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

/* Measurement, no counterpart in the reference: the vector-operation kernel's
 * launches timed with HIP events since creation (up to 8 per tick, recorded on
 * the ensemble's stream around each launch and read after the tick's wait),
 * their total device time [ms] and their algorithmic bytes (8 bytes per entry
 * of every vector each queued operation reads or writes). */
int  pihm_ens_op_stats(pihm_ensemble *e, long *launches, double *bytes, double *ms);

/* Autotuner of the launch shapes, choices the results do not depend on:
 * the placement of the batched RHS's element tiles over the XCDs (the
 * candidates of PIHM_TUNE_SPANS, pihm_ctx_set_tile_span) and the number of
 * 2,048-entry chunks each workgroup of the vector-operation kernel runs in
 * turn (1, 2, 4, 8).  It times reps batched RHS evaluations and reps
 * vector-operation launches (scale and linear sums) over the members that
 * are ready and not started, keeps the fastest of each, and reports them.
 * No member may be running; the members' states, carried river-edge fluxes
 * and NaN flags are left as they were.  pihm_ens_set_tuning sets the two
 * choices directly.  No counterpart in the reference. */
#define PIHM_ENS_TUNE_NG 4
typedef struct pihm_ens_tuning {
  int tile_span;
  int ops_chunks_per_wg;
  double rhs_ms[PIHM_TUNE_NSPAN];
  double ops_ms[PIHM_ENS_TUNE_NG];
} pihm_ens_tuning;
int  pihm_ens_autotune(pihm_ensemble *e, int reps, pihm_ens_tuning *out);
int  pihm_ens_set_tuning(pihm_ensemble *e, int tile_span, int ops_chunks_per_wg);

/* ---- serving: members that come and go ----------------------------------
 * An ensemble over a slot context (pihm_ctx_create_slots) starts with every
 * slot empty; slot k takes forcing of up to nrec_cap[k] records and
 * nstation_cap[k] stations.  pihm_ens_load_member puts a new solve into a
 * slot that is not running (empty, done or failed): mesh, forcing, initial
 * state (caller order) and the solver time t_end it runs to (a multiple of
 * cfg->stepsize); it starts running at the next pihm_ens_advance, beside
 * whatever else is running.  pihm_ens_advance runs at most max_ticks
 * synchronisations of all running members and returns after the first tick
 * in which any member finished or failed; it returns the number of such
 * members (0 if max_ticks ran out first), and *nfailed the failed ones.
 * A finished member's state stays readable (pihm_ens_get_state) until its
 * slot is loaded again. */
#define PIHM_ENS_EMPTY   0
#define PIHM_ENS_READY   1
#define PIHM_ENS_RUNNING 2
#define PIHM_ENS_DONE    3
#define PIHM_ENS_FAILED  (-1)
int  pihm_ens_create_slots(pihm_ctx *slots, const int32_t *nrec_cap, const int32_t *nstation_cap,
                           const pihm_run_config *cfg, pihm_ensemble **out);
int  pihm_ens_load_member(pihm_ensemble *e, int k, const pihm_mesh_desc *mesh, const pihm_forcing *forc,
                          const double *y0, double t_end);
int  pihm_ens_advance(pihm_ensemble *e, long max_ticks, int *nfailed);
int  pihm_ens_member_status(const pihm_ensemble *e, int k);

#ifdef __cplusplus
}
#endif
#endif
