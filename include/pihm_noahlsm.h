/*
 * pihm_noahlsm.h - the Noah land-surface step of the reference's _NOAH_ build
 * (Flux-PIHM), on the device.
 *
 * Every LSM step (src/pihm.c:25-63 with _NOAH_) the reference interpolates the
 * stations' meteorological records (ApplyMeteoForc, src/forcing.c:248-372;
 * IntrplForc :482-530) and runs Noah() (src/noah/noah.c:3-55) per element:
 * CalHum, FrozRain and SFlx (:166-609) with AlCalc, SnFrac, SnowNew, TDfCnd,
 * Snowz0, SfcDifOff, Penman, CanRes, NoPac or SnoPac (Evapo, DEvap, Transp,
 * PcpDrp), ShFlx (HRT, SnkSrc/FrH2O, TBnd, TmpAvg, HStep, Rosr12) and
 * SnowPack.  It leaves the net precipitation reaching the soil (wf.pcpdrp),
 * soil evaporation (wf.edir), transpiration (wf.ett) and its split over the
 * soil layers (wf.et[]), the soil temperatures and frozen water, and the snow
 * and canopy stores.
 *
 * pihm_noah_lsm_step does both on the device.  It works on the soil column of
 * a soil-water step (pihm_noahsoil.h) and on that step's context, whose
 * _NOAH_ branches must be on (pihm_ctx_set_noah): it reads and writes the
 * layer moisture smc/sh2o of the column, and writes pcpdrp/edir/ett, the
 * layer transpiration, sh2o[0] and GwTransp's share into the context for the
 * RHS, as the reference's next model steps read them.
 *
 * Configuration restated: rad_mode 0 (the station's SOLAR record is the
 * downward shortwave), no urban land cover, the bulk Richardson number of
 * SnoPac's frost adjustment at its initial 0 (nothing updates it in this
 * configuration).
 */
#ifndef PIHM_NOAHLSM_H
#define PIHM_NOAHLSM_H

#include <stdint.h>

#include "pihm_driver.h"
#include "pihm_noahsoil.h"
#include "pihm_rhs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Station records for Noah (METEO_TS blocks with _NOAH_): nrec records per
 * station at ascending times ftime [s], each with PRCP (kg m-2 s-1), SFCTMP
 * (K), RH (%), SFCSPD (m s-1), SOLAR (W m-2), LONGWV (W m-2), PRES (Pa). */
#define PIHM_NOAH_MET_NVAR 7
typedef struct pihm_noah_meteo {
  int32_t nstation, nrec;
  const double *ftime;   /* [nstation*nrec] */
  const double *data;    /* [nstation*nrec*PIHM_NOAH_MET_NVAR], record-major */
} pihm_noah_meteo;

/* Per element, caller order, as InitLc, InitSoil and InitLsm leave them
 * (src/init_lc.c, src/init_soil.c:24-74, src/noah/lsm_init.c): the root layer
 * ps.nroot (1-based, at most the column's nsoil), the element's station
 * (attrib.meteo_type - 1), lc.shdfac/cmcfactr/laimin/laimax/emissmin/
 * emissmax/albedomin/albedomax/z0min/z0max/snup/cfactr, epc.rgl/rsmin/rsmax/
 * topt/hs, soil.smcref/smcwlt/smcdry/quartz, ps.zlvl_wind; then the universal
 * values of the .lsm file with the calibration applied (soil.csoil, ps.sbeta,
 * salp, fxexp, czil, lvcoef, zbot, tbot) and ps.snoalb, ps.zlvl
 * (src/initialize.c:667-668).  The soil's smcmax, the layer thicknesses and
 * nsoil come from the soil-water step; smcmin, alpha and beta from the
 * context. */
typedef struct pihm_noah_lsm_desc {
  const int32_t *nroot, *station;
  const double *shdfac, *cmcfactr, *laimin, *laimax, *emissmin, *emissmax, *albedomin, *albedomax;
  const double *z0min, *z0max, *snup, *cfactr, *rgl, *rsmin, *rsmax, *topt, *hs;
  const double *smcref, *smcwlt, *smcdry, *quartz, *zlvl_wind;
  double csoil, sbeta, salp, fxexp, czil, lvcoef, zbot, tbot, snoalb, zlvl;
} pihm_noah_lsm_desc;

typedef struct pihm_noah_lsm pihm_noah_lsm;

/* Bind a Noah step to a soil-water step (and its context).  The per-element
 * states start as InitVar leaves them (src/initialize.c:660-700): no snow or
 * canopy water, albedo 0.18, emissivity 0.96, exchange coefficients 1e-4;
 * soil temperatures from pihm_noah_lsm_set_state. */
int  pihm_noah_lsm_create(pihm_noah_soil *soil, const pihm_noah_lsm_desc *desc, const pihm_noah_meteo *met,
                          pihm_noah_lsm **out);
void pihm_noah_lsm_destroy(pihm_noah_lsm *n);

/* Projected LAI per element (caller order) for the steps that follow (the
 * LAI forcing of ApplyLai, src/forcing.c:374-480). */
int  pihm_noah_lsm_set_lai(pihm_noah_lsm *n, const double *lai);

/* Soil temperatures stc [ne*11] (element-major, K), skin temperature t1, snow
 * water equivalent sneqv, snow depth snowh and canopy water cmc [ne] (m),
 * caller order; NULL leaves a field as it is. */
int  pihm_noah_lsm_set_state(pihm_noah_lsm *n, const double *stc, const double *t1, const double *sneqv,
                             const double *snowh, const double *cmc);

/* The carried state, caller order; any pointer may be NULL: stc [ne*11];
 * scalars [ne*PIHM_NOAH_NSCALAR] element-major in the order t1, sneqv, snowh,
 * sndens, cmc, ch, cm, snotime1, albedo, emissi. */
#define PIHM_NOAH_NSCALAR 10
int  pihm_noah_lsm_get_state(pihm_noah_lsm *n, double *stc, double *scalars);

/* One LSM step at time t [s] (on the records' clock) over dt [s]: station
 * interpolation, then Noah() per element; hands pcpdrp, edir, ett, et[],
 * sh2o[0] and GwTransp's share to the context.  Asynchronous on the
 * context's stream.  PIHM_ERR_ARG, where the reference exits, when t lies
 * outside a station's records. */
int  pihm_noah_lsm_step(pihm_noah_lsm *n, double t, double dt);

/* The last step's fluxes, caller order, [ne*PIHM_NOAH_NFLUX] element-major in
 * the order edir, ett, ec, pcpdrp, drip, esnow, etp, snomlt, ssoil, sheat,
 * eta, fdown, sncovr, beta, rc, pc (m s-1 for water, W m-2 for energy, as
 * the reference's wf and ef after Noah()), and et [ne*11] (m s-1); either
 * may be NULL.  Synchronises the stream. */
#define PIHM_NOAH_NFLUX 16
int  pihm_noah_lsm_get_fluxes(pihm_noah_lsm *n, double *flux, double *et);

/* Measurement: iters back-to-back steps at t on the context's stream, timed
 * with HIP events (the carried state advances).  No counterpart in the
 * reference. */
int  pihm_noah_lsm_time(pihm_noah_lsm *n, double t, double dt, int iters, double *ms_per_step);

/* The Flux-PIHM model-step loop on the device: every LSM step (cfg->etstep)
 * starts with pihm_noah_lsm_step at the step's time over etstep
 * (src/pihm.c:25-63), the RHS with the _NOAH_ branches runs in the
 * integrator, and every Summary runs MassBalance and the soil-water step
 * (src/pihm.c:95-110).  cfg->t_start is on the records' clock. */
int  pihm_run_gpu_flux(pihm_noah_lsm *n, const pihm_run_config *cfg, double *y_inout, pihm_run_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* PIHM_NOAHLSM_H */
