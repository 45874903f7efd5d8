/*
 * pihm_lsm.h - the land-surface step that produces the RHS's per-step
 * inputs, on the device.
 *
 * Every LSM step (src/pihm.c:25-70) the reference interpolates each
 * meteorological station's records to the current time (IntrplForc,
 * src/forcing.c:482-530; ApplyMeteoForc, src/forcing.c:305-344) and then,
 * per element, evaluates interception, snow and evapotranspiration
 * (IntcpSnowEt, src/is_sm_et.c:5-226).  That step writes the net
 * precipitation reaching the ground (pcpdrp), soil evaporation (edir) and
 * transpiration (ett) that ODE() reads, and carries the canopy storage (cmc)
 * and snow water equivalent (sneqv) to the next step.
 *
 * pihm_lsm_step does both on the device for a single-solve context and
 * writes pcpdrp/edir/ett straight into the context's step inputs: no upload
 * per step.  The vegetation inputs that change by month or by LAI forcing
 * (projected LAI, roughness length z0; MonthlyLai, MonthlyRl, MonthlyMf and
 * ApplyLai, src/forcing.c:374-798) and the snow melt factor are supplied by
 * the caller through pihm_lsm_set_veg when they change.
 */
#ifndef PIHM_LSM_H
#define PIHM_LSM_H

#include <stdint.h>
#include "pihm_bc.h"
#include "pihm_driver.h"
#include "pihm_nvector.h"
#include "pihm_rhs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-element parameters, caller order [nelem], as InitLc/InitSoil/InitTopo
 * leave them in elem_struct (src/init_lc.c:22-37, src/init_soil.c:24-74):
 * lc.albedomin/albedomax/shdfac/cmcfactr/cfactr, ps.zlvl_wind (the station's
 * WIND_LVL) and ps.rzd, epc.rgl/rsmin/rsmax/topt, soil.depth/porosity/
 * smcmin/smcwlt/smcref, topo.zmax; the element's meteorological station
 * (attrib.meteo_type - 1); calibration multipliers cal.edir/ec/ett. */
typedef struct pihm_lsm_desc {
  const double *albedomin, *albedomax, *shdfac, *cmcfactr, *cfactr;
  const double *zlvl_wind, *rzd;
  const double *rgl, *rsmin, *rsmax, *topt;
  const double *depth, *porosity, *smcmin, *smcwlt, *smcref, *zmax;
  const int32_t *station;
  double cal_edir, cal_ec, cal_ett;
} pihm_lsm_desc;

/* Meteorological records of every station (METEO_TS blocks of the .meteo
 * file): nrec records per station at ascending times ftime [s], each with
 * PRCP [kg m-2 s-1], SFCTMP [K], RH [%], SFCSPD [m s-1], SOLAR [W m-2]. */
#define PIHM_MET_NVAR 5
typedef struct pihm_meteo {
  int32_t nstation, nrec;
  const double *ftime;   /* [nstation*nrec] */
  const double *data;    /* [nstation*nrec*PIHM_MET_NVAR], record-major */
} pihm_meteo;

typedef struct pihm_lsm pihm_lsm;

/* Bind a land-surface step to a single-solve context (PIHM_ERR_ARG for a
 * batch).  Carried storages start at 0 (InitWState, src/initialize.c). */
int  pihm_lsm_create(pihm_ctx *ctx, const pihm_lsm_desc *desc, const pihm_meteo *met, pihm_lsm **out);
void pihm_lsm_destroy(pihm_lsm *lsm);

/* Projected LAI and roughness length per element (caller order) and the snow
 * melt factor [m s-1 K-1] for the steps that follow. */
int  pihm_lsm_set_veg(pihm_lsm *lsm, const double *lai, const double *z0, double meltf);

/* Carried canopy storage and snow water equivalent [m], caller order. */
int  pihm_lsm_set_state(pihm_lsm *lsm, const double *cmc, const double *sneqv);
int  pihm_lsm_get_state(pihm_lsm *lsm, double *cmc, double *sneqv);

/* One LSM step at time t [s] over stepsize [s]: interpolate the stations'
 * records, then IntcpSnowEt with the soil state of y (device vector, the
 * context's device order, as after Summary).  Writes pcpdrp/edir/ett into
 * the context's step inputs; asynchronous on the context's stream.  Returns
 * PIHM_ERR_ARG, where the reference exits, when t lies outside a station's
 * records (src/forcing.c:487-497). */
int  pihm_lsm_step(pihm_lsm *lsm, double t, double stepsize, N_Vector y);

/* Measurement: iters back-to-back steps at t on the context's stream, timed
 * with HIP events; *ms_per_step = elapsed / iters.  The carried storages
 * advance by iters steps.  No counterpart in the reference. */
int  pihm_lsm_time(pihm_lsm *lsm, double t, double stepsize, N_Vector y, int iters, double *ms_per_step);

/* The last step's pcpdrp, edir, ett, ec (canopy evaporation) and drip,
 * caller order; any pointer may be NULL.  Synchronises the stream. */
int  pihm_lsm_get_fluxes(pihm_lsm *lsm, double *pcpdrp, double *edir, double *ett, double *ec, double *drip);

/* A whole run with the device producers as the forcing: the model-step loop
 * of pihm_run_gpu (src/main.c:217-254, src/pihm.c:10-187) where every model
 * step starts with pihm_bc_apply at the step's time when bc is not NULL
 * (ApplyBc, src/pihm.c:17-24), and every LSM step (cfg->etstep) with
 * pihm_lsm_step at the step's time over etstep, as the reference's host path
 * calls IntcpSnowEt (src/pihm.c:60-63), on the state the last Summary saw.
 * cfg->t_start is on the records' clock.  The carried canopy and snow
 * storages stay in the LSM (pihm_lsm_get_state).  Returns PIHM_ERR_ARG when a
 * step's time lies outside the records or series. */
int  pihm_run_gpu_lsm(pihm_ctx *ctx, pihm_lsm *lsm, pihm_bc *bc, const pihm_run_config *cfg,
                      double *y_inout, pihm_run_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* PIHM_LSM_H */
