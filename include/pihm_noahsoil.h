/*
 * pihm_noahsoil.h - the soil-water step that follows every model step when the
 * reference is built with _NOAH_ (Flux-PIHM), on the device.
 *
 * Replaces, per element, Summary's water-table search and lateral-runoff split
 * (src/update.c:37-43: FindWaterTable src/noah/lsm_func.c:4-55, CalcLatFlx
 * :207-257) and NoahHydrol (src/noah/noah.c:57-94): the moisture clamps and
 * SmFlx (:1654-1712) with SRT (:2270-2425), WDfCnd (:2813-2861), SStep
 * (:2427-2481), Rosr12 (:1572-1620) and AdjSmProf (:3081-3172).  The reference
 * calls them after CVode returns at each model step (src/pihm.c:95-110).
 *
 * The step works on a context whose _NOAH_ branches are on
 * (pihm_ctx_set_noah): it reads the soil parameters the RHS reads (ksatv,
 * alpha, beta, smcmin), the groundwater of the context's state, the step's
 * edir (pihm_ctx_set_step_inputs) and the layer transpiration et of the last
 * pihm_ctx_set_noah_step, and leaves in the context what the next model
 * step's RHS reads: sh2o[0], fcr, nwtbl, and GwTransp's share recomputed from
 * the new water-table layer.  Single solves only.
 */
#ifndef PIHM_NOAHSOIL_H
#define PIHM_NOAHSOIL_H

#include <stdint.h>

#include "pihm_driver.h"
#include "pihm_rhs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pihm_noah_soil pihm_noah_soil;

/* Per element, caller order.  nsoil in [2, PIHM_NOAH_MAXLYR]; sldpth holds
 * PIHM_NOAH_MAXLYR entries per element (element-major, as ps.sldpth[MAXLYR]),
 * the first nsoil used; smcmax - smcmin must equal the mesh's porosity
 * (src/init_soil.c:34-36); frzx is ps.frzx (src/noah/lsm_init.c:43). */
typedef struct pihm_noah_soil_desc {
  const int32_t *nsoil;
  const double *sldpth;
  const double *smcmax;
  const double *frzx;
} pihm_noah_soil_desc;

int  pihm_noah_soil_create(pihm_ctx *ctx, const pihm_noah_soil_desc *desc, pihm_noah_soil **out);
void pihm_noah_soil_destroy(pihm_noah_soil *s);
/* Layer moisture smc and liquid water sh2o (element-major, PIHM_NOAH_MAXLYR
 * per element, m3 m-3) and optionally fcr; also copies sh2o[0] and fcr into
 * the context for the RHS.  The carried layer fluxes and runoff3 start at 0. */
int  pihm_noah_soil_set_state(pihm_noah_soil *s, const double *smc, const double *sh2o, const double *fcr);
/* Any pointer may be NULL.  Per-layer arrays element-major as above. */
int  pihm_noah_soil_get_state(pihm_noah_soil *s, double *smc, double *sh2o, double *fcr, int32_t *nwtbl,
                              double *satdpth, double *runoff2_lyr, double *smflxv, double *runoff3);
/* One step of length dt (ctrl.stepsize) at the state y_dev (the context's
 * device state, as after CVode's return).  infil and runoff2 are MassBalance's
 * re-derived infiltration and subsurface runoff of the step (src/update.c:
 * 107-160; m s-1, caller order), which Summary hands to CalcLatFlx and SRT. */
int  pihm_noah_soil_step(pihm_noah_soil *s, double dt, const double *y_dev, const double *infil,
                         const double *runoff2);
/* The same step with its inputs taken on the device, as Summary does
 * (src/update.c:3-60, MassBalance :107-160): the fluxes are evaluated at y_dev
 * (the carried river-edge flux is restored afterwards), the infiltration and
 * subsurface runoff are re-derived from the storage change since the previous
 * call, with the _NOAH_ EtExtract split, and the soil-water step follows.
 * pihm_noah_soil_start records the initial storage (the first Summary). */
int  pihm_noah_soil_start(pihm_noah_soil *s, const double *y_dev);
int  pihm_noah_soil_summary(pihm_noah_soil *s, double dt, const double *y_dev);
/* Runs pihm_noah_soil_start at the driver's first Summary and
 * pihm_noah_soil_summary (with ctrl.stepsize) at every later one; NULL detaches. */
int  pihm_gpu_backend_set_noah_soil(pihm_gpu_backend *b, pihm_noah_soil *s);
/* pihm_run_gpu with the soil-water step at every Summary (the context's
 * _NOAH_ branches on); the land-surface values come from the forcing. */
int  pihm_run_gpu_noah(pihm_ctx *ctx, const pihm_forcing *forc, pihm_noah_soil *soil,
                       const pihm_run_config *cfg, double *y_inout, pihm_run_stats *stats);
/* HIP-event time of iters back-to-back steps (kernel only), ms per step. */
int  pihm_noah_soil_time(pihm_noah_soil *s, double dt, const double *y_dev, int iters, double *ms);

#ifdef __cplusplus
}
#endif

#endif
