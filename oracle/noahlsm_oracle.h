/*
 * noahlsm_oracle.h - CPU restatement of the Noah land-surface step (Noah() and
 * SFlx, src/noah/noah.c:3-609).  TEST INFRASTRUCTURE ONLY (see pihm_oracle.h).
 */
#ifndef NOAHLSM_ORACLE_H
#define NOAHLSM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* What one element reads that does not change during a run: lc, soil,
 * epc and the ps fields InitLsm/InitLc/InitSoil set (src/noah/lsm_init.c,
 * src/init_lc.c, src/init_soil.c, src/initialize.c:660-680). */
typedef struct oracle_noahlsm_par {
    int nsoil, nroot;
    double sldpth[11];
    double shdfac, cmcfactr, laimin, laimax, emissmin, emissmax, albedomin, albedomax, z0min, z0max;
    double snup, cfactr, rgl, rsmin, rsmax, topt, hs;
    double smcmax, smcmin, smcref, smcwlt, smcdry, quartz, csoil, alpha, beta;
    double sbeta, salp, fxexp, czil, lvcoef, zbot, tbot, snoalb, zlvl, zlvl_wind;
} oracle_noahlsm_par;

/* Carried from one LSM step to the next (ws, es, and the ps fields the
 * step reads before it writes them). */
typedef struct oracle_noahlsm_state {
    double smc[11], sh2o[11], stc[11];
    double t1, sneqv, snowh, sndens, cmc, ch, cm, snotime1, albedo, emissi, q1;
} oracle_noahlsm_state;

/* The element's forcing at the step (ApplyMeteoForc with _NOAH_,
 * src/forcing.c:340-353; prcp in m s-1) and its projected LAI. */
typedef struct oracle_noahlsm_forc {
    double sfctmp, rh, sfcspd, soldn, longwave, sfcprs, prcp, proj_lai;
} oracle_noahlsm_forc;

/* What the step leaves: wf.edir/ett/ec after Noah()'s unit round trip, the
 * per-layer transpiration wf.et, wf.pcpdrp, and diagnostics. */
typedef struct oracle_noahlsm_out {
    double edir, ett, ec, et[11], pcpdrp, drip, esnow, etp, snomlt, ssoil, sheat, eta, fdown, sncovr, beta;
    double rc, pc, flx1, flx2, flx3, dew, soilw, soilm, cmcmax, rr, rch, epsca, q2, q2sat, th2;
} oracle_noahlsm_out;

void oracle_noah_element(const oracle_noahlsm_par *p, oracle_noahlsm_state *st, const oracle_noahlsm_forc *f,
                         double dt, oracle_noahlsm_out *o);
void oracle_noah_step(int ne, double dt, const oracle_noahlsm_par *par, oracle_noahlsm_state *st,
                      const oracle_noahlsm_forc *forc, oracle_noahlsm_out *out);

/* pieces pinned on their own */
double oracle_CSnow(double dsnow);
double oracle_SnFrac(double sneqv, double snup, double salp);
double oracle_TDfCnd(double smc, double qz, double smcmax, double smcmin, double sh2o);
double oracle_Pspmu(double xx);
double oracle_Psphu(double xx);
void   oracle_SfcDifOff(double zlvl, double zlvl_wind, double z0, double czil, double sfcspd, double t1v,
                        double th2v, double *cm, double *ch);
double oracle_TBnd(double tu, double tb, const double *zsoil, double zbot, int k, int nsoil);
double oracle_TmpAvg(double tup, double tm, double tdn, const double *zsoil, int k);
double oracle_FrH2O(double tkelv, double smc, double sh2o, double smcmax, double smcmin, double alpha,
                    double beta);
void   oracle_SnowPack(double esd, double dtsec, double *snowh, double *sndens, double tsnow, double tsoil);

#ifdef __cplusplus
}
#endif
#endif
