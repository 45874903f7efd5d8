/*
 * pihm_output.h - model output on the device: the reference's print
 * variables, averaged between prints, written in its file formats.
 *
 * The reference keeps, per print variable, a running sum of the variable's
 * values at each update (UpdPrintVar, src/print.c:190-211: after every model
 * step for hydrologic variables, after every LSM step for land-surface ones)
 * and, when PrintNow says so (src/print.c:814-856: yearly, monthly, daily,
 * hourly, or every intvl seconds of simulated time), writes the average
 * (sum / count) of every element or segment in the reference's order
 * (PrintData, src/print.c:545-603):
 *   <prefix>.<ext>.dat  binary: the time (double, seconds since the epoch)
 *                       then one double per element or segment, per print;
 *   <prefix>.<ext>.txt  with ascii: "YYYY-MM-DD HH:MM" then "\t%lf" per
 *                       value, one line per print.
 * Here the sums stay in HBM and only the printed averages cross to the host.
 * The flux variables (recharge, surfflx0-2, subflx0-2, rivflx0-10) are the
 * fluxes of one RHS evaluation at the model step's final state, made for the
 * update with the carried river-edge flux left unchanged; the reference
 * prints whatever its integrator's last RHS call left in elem.wf.  infil is
 * MassBalance's (src/update.c:107-160), which Summary writes over the RHS's:
 * the storage change of unsat + gw since the previous model step times the
 * porosity over the step, plus the subsurface outflow and the soil's share of
 * evaporation and transpiration of that evaluation, floored at 0.
 * pihm_output_write_ic writes the restart file of src/print.c:250-300: per
 * element cmc, sneqv, surf, unsat, gw, then per segment stage, gw (doubles).
 */
#ifndef PIHM_OUTPUT_H
#define PIHM_OUTPUT_H

#include <stdint.h>
#include "pihm_lsm.h"
#include "pihm_nvector.h"
#include "pihm_rhs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Print variables and their file extensions (src/map_output.c:30-255). */
enum {
  PIHM_OUT_SURF = 0,     /* surf      elem  model step */
  PIHM_OUT_UNSAT,        /* unsat     elem  model step */
  PIHM_OUT_GW,           /* gw        elem  model step */
  PIHM_OUT_RIVSTG,       /* stage     river model step */
  PIHM_OUT_RIVGW,        /* rivgw     river model step */
  PIHM_OUT_SNOW,         /* snow      elem  LSM step   */
  PIHM_OUT_CMC,          /* is        elem  LSM step   */
  PIHM_OUT_INFIL,        /* infil     elem  model step */
  PIHM_OUT_RECHARGE,     /* recharge  elem  model step */
  PIHM_OUT_EC,           /* ec        elem  LSM step   */
  PIHM_OUT_ETT,          /* ett       elem  LSM step   */
  PIHM_OUT_EDIR,         /* edir      elem  LSM step   */
  PIHM_OUT_RIVFLX0,      /* rivflx0 .. rivflx10: river model step (rivflow[k]) */
  PIHM_OUT_SUBFLX0 = PIHM_OUT_RIVFLX0 + 11,   /* subflx0 .. subflx2: elem (subsurf[k]) */
  PIHM_OUT_SURFFLX0 = PIHM_OUT_SUBFLX0 + 3,   /* surfflx0 .. surfflx2: elem (ovlflow[k]) */
  PIHM_OUT_NKIND = PIHM_OUT_SURFFLX0 + 3
};
/* Print intervals: these codes (src/include/pihm_const.h:231-234) or a
 * positive number of seconds. */
#define PIHM_OUT_YEARLY  (-1)
#define PIHM_OUT_MONTHLY (-2)
#define PIHM_OUT_DAILY   (-3)
#define PIHM_OUT_HOURLY  (-4)
#define PIHM_OUT_HYDROL_STEP 0
#define PIHM_OUT_LS_STEP     1

typedef struct pihm_output_var {
  int32_t kind;
  int32_t intvl;
} pihm_output_var;

typedef struct pihm_output pihm_output;

/* Files <prefix>.<ext>.dat (and .txt with ascii) are created for every
 * variable.  lsm may be NULL when no land-surface variable is asked for.
 * Single-solve contexts only. */
int  pihm_output_create(pihm_ctx *ctx, pihm_lsm *lsm, const char *prefix, const pihm_output_var *vars,
                        int nvars, int ascii, pihm_output **out);
void pihm_output_destroy(pihm_output *out);

/* The initial state (ws0 after initialisation), needed by infil before the
 * first model-step update; pihm_run_gpu_producers calls it. */
int  pihm_output_start(pihm_output *out, N_Vector y0);

/* UpdPrintVar: add the current values of the variables of this update step
 * (PIHM_OUT_HYDROL_STEP with the model step's final state y, or
 * PIHM_OUT_LS_STEP after a land-surface step) to their sums. */
int  pihm_output_update(pihm_output *out, int module_step, N_Vector y);

/* PrintData at time t [s since the epoch], lapse = t - start time [s]. */
int  pihm_output_print(pihm_output *out, double t, double lapse);

/* The restart file at path: state y and the LSM's storages (0 without one). */
int  pihm_output_write_ic(pihm_output *out, const char *path, N_Vector y);

/* A whole run on the device with every producer and the output: the
 * sequence of PIHM() (src/pihm.c:10-187) for each model step -- ApplyBc (bc),
 * the LSM step and its output update every LSM step (lsm), SolveCVode,
 * Summary, the model-step output update and print (out).  Any of lsm, bc,
 * out may be NULL (no LSM: the step inputs stay as set). */
typedef struct pihm_producers {
  pihm_lsm *lsm;
  pihm_bc *bc;
  pihm_output *out;
} pihm_producers;
int  pihm_run_gpu_producers(pihm_ctx *ctx, const pihm_producers *p, const pihm_run_config *cfg,
                            double *y_inout, pihm_run_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* PIHM_OUTPUT_H */
