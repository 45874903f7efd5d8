/*
 * lsm_oracle.c - CPU restatement of the reference's per-step producers: the simple
 * land-surface step and the boundary-condition series.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/ may load this code, as the checker
 * of the device land-surface step.  The product library never links it.
 *
 *   oracle_meteo_interp   IntrplForc (src/forcing.c:482-530) for every
 *                         station's records
 *   oracle_bc_apply       ApplyElemBc / ApplyRiverBc (src/forcing.c:127-244)
 *   oracle_intcp_snow_et  IntcpSnowEt (src/is_sm_et.c:5-226) with
 *                         ApplyMeteoForc's per-element values
 *                         (src/forcing.c:340-349)
 *
 * Operation order as in the reference; libm's exp, log, pow and cos.
 */
#include <math.h>

#include "../include/pihm_bc.h"
#include "../include/pihm_lsm.h"

/* src/include/pihm_const.h:9-15 */
#define O_CP 1004.0
#define O_LVH2O 2.501e6
#define O_SIGMA 5.67e-8
#define O_RD 287.04
#define O_RV 461.5

/* IntrplForc (src/forcing.c:482-530) on one series: ft[nrec], d[nrec*nvar]
 * record-major, into out[v * ostride]; 1 when t is outside (the reference
 * exits). */
static int interp_series(const double *ft, const double *d, int nrec, int nvar, double t, double *out,
                         long ostride)
{
    int first = 1, last = nrec - 1, middle, j;
    if (t < ft[0] || t > ft[nrec - 1]) return 1;
    while (first <= last) {
        middle = (first + last) / 2;
        if (t >= ft[middle - 1] && t <= ft[middle]) {
            for (j = 0; j < nvar; j++)
                out[j * ostride] =
                    ((ft[middle] - t) * d[(long)(middle - 1) * nvar + j] +
                     (t - ft[middle - 1]) * d[(long)middle * nvar + j]) /
                    (ft[middle] - ft[middle - 1]);
            break;
        } else if (ft[middle] > t) {
            last = middle - 1;
        } else {
            first = middle + 1;
        }
    }
    return 0;
}

/* Each station's records at time t into met[v * nstation + k]; 1 when t is
 * outside a station's records. */
int oracle_meteo_interp(const pihm_meteo *m, double t, double *met)
{
    int k;
    for (k = 0; k < m->nstation; k++)
        if (interp_series(m->ftime + (long)k * m->nrec, m->data + (long)k * m->nrec * PIHM_MET_NVAR,
                          m->nrec, PIHM_MET_NVAR, t, met + k, m->nstation))
            return 1;
    return 0;
}

/* The same for records of nvar variables: met[v * nstation + k]. */
int oracle_meteo_interp_n(const double *ftime, const double *data, int nstation, int nrec, int nvar, double t,
                          double *met)
{
    int k;
    for (k = 0; k < nstation; k++)
        if (interp_series(ftime + (long)k * nrec, data + (long)k * nrec * nvar, nrec, nvar, t, met + k, nstation))
            return 1;
    return 0;
}

/* ApplyElemBc and ApplyRiverBc (src/forcing.c:127-177, 214-244): every series
 * interpolated at t, then each element edge (elem_bc[3*i+j], caller order)
 * and river segment with a nonzero code takes series |code|-1's value.
 * Either set may be absent (NULL).  1 when t is outside a series. */
int oracle_bc_apply(const pihm_ts *ets, const int32_t *etype, int ne, double *elem_bc,
                    const pihm_ts *rts, const int32_t *rtype, int nr, double *river_bc, double t)
{
    double ev[4096], rv[4096];
    int k, i, j;
    const int nts_e = ets ? ets->nts : 0, nts_r = rts ? rts->nts : 0;
    if (nts_e > 4096 || nts_r > 4096) return 2;
    for (k = 0; k < nts_e; k++)
        if (interp_series(ets->ftime + (long)k * ets->nrec, ets->value + (long)k * ets->nrec, ets->nrec, 1,
                          t, ev + k, 1))
            return 1;
    for (k = 0; k < nts_r; k++)
        if (interp_series(rts->ftime + (long)k * rts->nrec, rts->value + (long)k * rts->nrec, rts->nrec, 1,
                          t, rv + k, 1))
            return 1;
    for (i = 0; etype && i < ne; i++)
        for (j = 0; j < 3; j++) {
            const int c = etype[3 * i + j];
            if (c > 0) elem_bc[3 * i + j] = ev[c - 1];
            else if (c < 0) elem_bc[3 * i + j] = ev[-c - 1];
        }
    for (i = 0; rtype && i < nr; i++) {
        const int c = rtype[i];
        if (c > 0) river_bc[i] = rv[c - 1];
        else if (c < 0) river_bc[i] = rv[-c - 1];
    }
    return 0;
}

/* One LSM step for ne elements (caller order).  unsat/gw: the soil state as
 * after Summary.  cmc/sneqv are updated in place; pcpdrp/edir/ett/ec/drip
 * receive the step's fluxes. */
void oracle_intcp_snow_et(int ne, const pihm_lsm_desc *d, const double *met, int nstation,
                          const double *lai_e, const double *z0_e, double meltf,
                          const double *unsat_e, const double *gw_e, double stepsize,
                          double *cmc_e, double *sneqv_e, double *pcpdrp_e, double *edir_e,
                          double *ett_e, double *ec_e, double *drip_e)
{
    const double TSNOW = -3.0, TRAIN = 1.0, T0 = 0.0;
    int i;
    for (i = 0; i < ne; i++) {
        const int st = d->station[i];
        double prcp, sfctmp, wind, rh, soldn, vp, pres, qv, qvsat, lai, albedo, radnet;
        double frac_snow, snow_rate, melt_rate, intcp_max, z0, ra, gamma, delta, etp;
        double satn, betas, edir, ec, ett, drip, isval, cmc, sneqv;
        double fr, alphar, etas, gammas, rs, pc;
        const double unsat = unsat_e[i], gw = gw_e[i];

        /* ApplyMeteoForc: src/forcing.c:340-349 */
        prcp = met[0 * nstation + st] / 1000.0;
        soldn = met[4 * nstation + st];
        soldn = (soldn > 0.0) ? soldn : 0.0;

        /* src/is_sm_et.c:48-67 */
        albedo = 0.5 * (d->albedomin[i] + d->albedomax[i]);
        radnet = soldn * (1.0 - albedo);
        sfctmp = met[1 * nstation + st] - 273.15;
        wind = met[3 * nstation + st];
        rh = met[2 * nstation + st] / 100.0;
        vp = 611.2 * exp(17.67 * sfctmp / (sfctmp + 243.5)) * rh;
        pres = 101.325 * 1.0e3 * pow((293.0 - 0.0065 * d->zmax[i]) / 293.0, 5.26);
        qv = 0.622 * vp / pres;
        qvsat = 0.622 * (vp / rh) / pres;
        lai = lai_e[i];

        /* snow: src/is_sm_et.c:70-88 */
        cmc = cmc_e[i];
        sneqv = sneqv_e[i];
        frac_snow = (sfctmp < TSNOW) ? 1.0 :
            ((sfctmp > TRAIN) ? 0.0 : (TRAIN - sfctmp) / (TRAIN - TSNOW));
        snow_rate = frac_snow * prcp;
        sneqv += snow_rate * stepsize;
        melt_rate = (sfctmp > T0) ? (sfctmp - T0) * meltf : 0.0;
        if (sneqv > melt_rate * stepsize) {
            sneqv -= melt_rate * stepsize;
        } else {
            melt_rate = sneqv / stepsize;
            sneqv = 0.0;
        }

        /* potential evaporation: src/is_sm_et.c:90-104 */
        intcp_max = d->cmcfactr[i] * lai * d->shdfac[i];
        z0 = z0_e[i];
        ra = log(d->zlvl_wind[i] / z0) * log(10.0 * d->zlvl_wind[i] / z0) / (wind * 0.16);
        gamma = 4.0 * 0.7 * O_SIGMA * O_RD / O_CP * pow(sfctmp + 273.15, 4) / (pres / ra) + 1.0;
        delta = O_LVH2O * O_LVH2O * 0.622 / O_RV / O_CP / pow(sfctmp + 273.15, 2) * qvsat;
        etp = (radnet * delta + gamma * (1.2 * O_LVH2O * (qvsat - qv) / ra)) /
            (1000.0 * O_LVH2O * (delta + gamma));

        /* soil evaporation: src/is_sm_et.c:106-127 */
        if (d->depth[i] - gw < d->rzd[i]) {
            satn = 1.0;
        } else {
            const double x = unsat / (d->depth[i] - gw);
            satn = (x > 1.0) ? 1.0 : ((x < 0.0) ? 0.0 : 0.5 * (1.0 - cos(3.14 * x)));
        }
        betas = (satn * d->porosity[i] + d->smcmin[i] - d->smcwlt[i]) /
            (d->smcref[i] - d->smcwlt[i]);
        betas = (betas < 0.0001) ? 0.0001 : ((betas > 1.0) ? 1.0 : betas);
        edir = (1.0 - d->shdfac[i]) * pow(betas, 2) * etp;
        edir *= d->cal_edir;
        edir = (edir < 0.0) ? 0.0 : edir;

        /* canopy evaporation, transpiration, drip: src/is_sm_et.c:129-175 */
        if (lai > 0.0) {
            const double wet = ((cmc < 0.0) ? 0.0 : ((cmc > intcp_max) ? intcp_max : cmc)) / intcp_max;
            ec = d->shdfac[i] * pow(wet, d->cfactr[i]) * etp;
            ec *= d->cal_ec;
            ec = (ec < 0.0) ? 0.0 : ec;
            fr = 1.1 * radnet / (d->rgl[i] * lai);
            fr = (fr < 0.0) ? 0.0 : fr;
            alphar = (1.0 + fr) / (fr + (d->rsmin[i] / d->rsmax[i]));
            alphar = (alphar > 10000.0) ? 10000.0 : alphar;
            etas = 1.0 - 0.0016 * (pow((d->topt[i] - 273.15 - sfctmp), 2));
            etas = (etas < 0.0001) ? 0.0001 : etas;
            gammas = 1.0 / (1.0 + 0.00025 * (vp / rh - vp));
            gammas = (gammas < 0.01) ? 0.01 : gammas;
            rs = d->rsmin[i] * alphar / (betas * lai * etas * gammas);
            rs = (rs > d->rsmax[i]) ? d->rsmax[i] : rs;
            pc = (1.0 + delta / gamma) / (1.0 + rs / ra + delta / gamma);
            /* (cmc < 0 ? 0 : clamp(cmc) / intcp_max): the reference's
             * precedence, src/is_sm_et.c:160-163 */
            ett = d->shdfac[i] * pc *
                (1.0 - pow((cmc < 0.0) ? 0.0 :
                           ((cmc > intcp_max) ? intcp_max : cmc) / intcp_max, d->cfactr[i])) * etp;
            ett *= d->cal_ett;
            ett = (ett < 0.0) ? 0.0 : ett;
            ett = ((gw < (d->depth[i] - d->rzd[i])) && unsat <= 0.0) ? 0.0 : ett;
            drip = (cmc <= 0.0) ? 0.0 : 6.52E-7 * intcp_max * exp(3.89 * cmc / intcp_max);
        } else {
            ett = 0.0;
            ec = 0.0;
            drip = 0.0;
        }
        if (drip < 0.0) drip = 0.0;
        if (drip * stepsize > cmc) drip = cmc / stepsize;

        /* interception storage: src/is_sm_et.c:187-224 */
        isval = cmc + (1.0 - frac_snow) * prcp * d->shdfac[i] * stepsize -
            ec * stepsize - drip * stepsize;
        if (isval > intcp_max) {
            cmc = intcp_max;
            drip += (isval - intcp_max) / stepsize;
        } else if (isval < 0.0) {
            cmc = 0.0;
            if (ec + drip > 0.0) {
                ec = ec / (ec + drip) *
                    (cmc + (1.0 - frac_snow) * prcp * d->shdfac[i] * stepsize);
                drip = drip / (ec + drip) *
                    (cmc + (1.0 - frac_snow) * prcp * d->shdfac[i] * stepsize);
            }
        } else {
            cmc = isval;
        }

        cmc_e[i] = cmc;
        sneqv_e[i] = sneqv;
        pcpdrp_e[i] = (1.0 - d->shdfac[i]) * (1.0 - frac_snow) * prcp + drip + melt_rate;
        edir_e[i] = edir;
        ett_e[i] = ett;
        ec_e[i] = ec;
        drip_e[i] = drip;
    }
}
