/*
 * noahsoil_oracle.c - CPU restatement of the Flux-PIHM soil-water step that
 * follows every model step when the reference is built with _NOAH_.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/ load this code, as the checker of the
 * device soil-water step.  The product library never links it.
 *
 * Per element, in the reference's order and operation order:
 *   Summary (src/update.c:37-43): FindWaterTable (src/noah/lsm_func.c:4-55)
 *     and CalcLatFlx (src/noah/lsm_func.c:207-257), which spreads the
 *     subsurface runoff over the saturated layers;
 *   NoahHydrol (src/noah/noah.c:57-94): FindWaterTable again, the moisture
 *     clamps, and SmFlx (:1654-1712): SRT (:2270-2425, frozen-soil factor and
 *     the tridiagonal coefficients, WDfCnd :2813-2861), SStep (:2427-2481,
 *     Rosr12 :1572-1620, AdjSmProf :3081-3172, layer fluxes).
 * The layer arrays of one element are its own (nsoil <= 11 entries).
 */
#include <math.h>
#include <string.h>

#include "pihm_oracle.h"

#define NS_MAXLYR 11
#define SH2OMIN 0.02            /* src/include/pihm_const.h:81 */

/* FindWaterTable: src/noah/lsm_func.c:4-55 */
int oracle_FindWaterTable(const double *sldpth, int nsoil, double gw, double *satdpth)
{
    int layer = -999, j;
    double dsum = 0.0, depth = 0.0;
    for (j = 0; j < NS_MAXLYR; j++) satdpth[j] = 0.0;
    for (j = 0; j < nsoil; j++) depth += sldpth[j];
    if (gw <= 0.0) {
        layer = nsoil;
        satdpth[nsoil - 1] = 1.0E-3;
    } else if (gw > depth) {
        layer = 0;
        for (j = 0; j < nsoil; j++) satdpth[j] = sldpth[j];
    } else {
        for (j = nsoil - 1; j >= 0; j--) {
            if (dsum + sldpth[j] > gw) {
                satdpth[j] = gw - dsum;
                layer = j + 1;
                break;
            }
            satdpth[j] = sldpth[j];
            dsum += sldpth[j];
        }
    }
    return layer;
}

/* CalcLatFlx: src/noah/lsm_func.c:207-257 (without _CYCLES_) */
void oracle_CalcLatFlx(int nsoil, const double *satdpth, double runoff2, double *runoff2_lyr)
{
    double sattot = 0.0;
    int ks;
    for (ks = 0; ks < nsoil; ks++) sattot += satdpth[ks];
    if (sattot <= 0.0) {
        runoff2_lyr[nsoil - 1] = runoff2;
    } else {
        for (ks = 0; ks < nsoil; ks++) runoff2_lyr[ks] = satdpth[ks] / sattot * runoff2;
    }
}

/* WDfCnd: src/noah/noah.c:2813-2861 (van Genuchten form of Flux-PIHM) */
void oracle_WDfCnd(double *wdf, double *wcnd, double smc, double sicemax, double smcmax, double smcmin,
                   double ksatv, double alpha, double beta)
{
    double expon, factr1, factr2, vkwgt, satkfunc, dpsidsm;
    factr1 = 0.05 / (smcmax - smcmin);
    factr2 = (smc - smcmin) / (smcmax - smcmin);
    factr2 = (factr2 > 1.0 - 5.0e-4) ? 1.0 - 5.0e-4 : factr2;
    factr2 = (factr2 < 0.0 + 5.0e-4) ? 5.0e-4 : factr2;
    factr1 = (factr1 < factr2) ? factr1 : factr2;
    expon = 1.0 - 1.0 / beta;
    satkfunc = oracle_KrFunc(beta, factr2);
    dpsidsm = (1.0 - expon) / alpha / expon / (smcmax - smcmin) *
        pow(pow(factr2, -1.0 / expon) - 1.0, -expon) * pow(factr2, -(1.0 / expon + 1.0));
    *wcnd = ksatv * satkfunc;
    *wdf = *wcnd * dpsidsm;
    if (sicemax > 0.0) {
        vkwgt = 1.0 / (1.0 + pow(500.0 * sicemax, 3.0));
        satkfunc = oracle_KrFunc(beta, factr1);
        dpsidsm = (1.0 - expon) / alpha / expon / (smcmax - smcmin) *
            pow(pow(factr1, -1.0 / expon) - 1.0, -expon) * pow(factr1, -(1.0 / expon + 1.0));
        *wdf = vkwgt * *wdf + (1.0 - vkwgt) * dpsidsm * satkfunc * ksatv;
    }
}

/* Rosr12: src/noah/noah.c:1572-1620, the tridiagonal solve a x[k-1] + b x[k] + c x[k+1] = d */
void oracle_Rosr12(double *p, const double *a, const double *b, double *c, const double *d, double *delta,
                   int nsoil)
{
    int k, kk;
    c[nsoil - 1] = 0.0;
    p[0] = -c[0] / b[0];
    delta[0] = d[0] / b[0];
    for (k = 1; k < nsoil; k++) {
        p[k] = -c[k] * (1.0 / (b[k] + a[k] * p[k - 1]));
        delta[k] = (d[k] - a[k] * delta[k - 1]) * (1.0 / (b[k] + a[k] * p[k - 1]));
    }
    p[nsoil - 1] = delta[nsoil - 1];
    for (k = 1; k < nsoil; k++) {
        kk = nsoil - k - 1;
        p[kk] = p[kk] * p[kk + 1] + delta[kk];
    }
}

/* SRT: src/noah/noah.c:2270-2425 (without _CYCLES_) */
static void SRT(const oracle_noah_col *in, const double *zsoil, oracle_noah_state *st, double *rhstt,
                const double *sice, double *ai, double *bi, double *ci)
{
    int j, jj, k, ks, nsoil = in->nsoil;
    double ddz, ddz2, denom, denom2, numer, pddum, mxsmc, mxsmc2, sicemax, wcnd, wcnd2, wdf, wdf2;
    double dsmdz, dsmdz2, dice, acrt, sum, ialp1;
    const double CVFRZ = 3.0;
    sicemax = 0.0;
    for (ks = 0; ks < nsoil; ks++) sicemax = (sice[ks] > sicemax) ? sice[ks] : sicemax;
    dice = -zsoil[0] * sice[0];
    for (ks = 1; ks < nsoil; ks++) dice += (zsoil[ks - 1] - zsoil[ks]) * sice[ks];
    st->fcr = 1.0;
    if (dice > 1.0e-2) {
        acrt = CVFRZ * in->frzx / dice;
        sum = 1.0;
        ialp1 = (int)round(CVFRZ) - 1;
        for (j = 1; j < ialp1 + 1; j++) {
            k = 1;
            for (jj = j + 1; jj < ialp1; jj++) k *= jj;
            sum += pow(acrt, CVFRZ - (double)j) / (double)k;
        }
        st->fcr = 1.0 - exp(-acrt) * sum;
    }
    pddum = in->infil;
    mxsmc = st->sh2o[0];
    dsmdz = (st->sh2o[0] - st->sh2o[1]) / (-0.5 * zsoil[1]);
    oracle_WDfCnd(&wdf, &wcnd, mxsmc, sicemax, in->smcmax, in->smcmin, in->ksatv, in->alpha, in->beta);
    ddz = 1.0 / (-0.5 * zsoil[1]);
    ai[0] = 0.0;
    bi[0] = wdf * ddz / (-zsoil[0]);
    ci[0] = -bi[0];
    rhstt[0] = (wdf * dsmdz + wcnd - pddum + in->edir + in->et[0]) / zsoil[0];
    rhstt[0] += st->runoff2_lyr[0] / zsoil[0];
    ddz2 = 0.0;
    for (k = 1; k < nsoil; k++) {
        denom2 = (zsoil[k - 1] - zsoil[k]);
        if (k < nsoil - 1) {
            mxsmc2 = st->sh2o[k];
            denom = zsoil[k - 1] - zsoil[k + 1];
            dsmdz2 = (st->sh2o[k] - st->sh2o[k + 1]) / (denom * 0.5);
            oracle_WDfCnd(&wdf2, &wcnd2, mxsmc2, sicemax, in->smcmax, in->smcmin, in->ksatv, in->alpha, in->beta);
            ddz2 = 2.0 / denom;
            ci[k] = -wdf2 * ddz2 / denom2;
        } else {
            wdf2 = 0.0;
            wcnd2 = 0.0;
            dsmdz2 = 0.0;
            ci[k] = 0.0;
        }
        numer = (wdf2 * dsmdz2) + wcnd2 - (wdf * dsmdz) - wcnd + in->et[k];
        numer += st->runoff2_lyr[k];
        rhstt[k] = numer / (-denom2);
        ai[k] = -wdf * ddz / denom2;
        bi[k] = -(ai[k] + ci[k]);
        if (k != nsoil - 1) {
            wdf = wdf2;
            wcnd = wcnd2;
            dsmdz = dsmdz2;
            ddz = ddz2;
        }
    }
}

/* AdjSmProf: src/noah/noah.c:3081-3172 */
static void AdjSmProf(const oracle_noah_col *in, const double *zsoil, const double *sice, double dt,
                      oracle_noah_state *st)
{
    int k, nsoil = in->nsoil;
    double ddz, sh2omid[NS_MAXLYR], wplus, stot, stotmin;
    const double smcmax = in->smcmax, smcmin = in->smcmin;
    wplus = 0.0;
    st->runoff3 = 0.0;
    for (k = nsoil - 1; k >= 0; k--) {
        ddz = (k != 0) ? zsoil[k - 1] - zsoil[k] : -zsoil[0];
        sh2omid[k] = st->sh2o[k] + wplus / ddz;
        stot = sh2omid[k] + sice[k];
        if (stot > smcmax) {
            st->smc[k] = smcmax;
            wplus = (stot - smcmax) * ddz;
        } else {
            stotmin = (st->satdpth[k] * smcmax + (in->sldpth[k] - st->satdpth[k]) * (smcmin + SH2OMIN)) /
                in->sldpth[k];
            stotmin = (stotmin > smcmax) ? smcmax : stotmin;
            stotmin = (stotmin < smcmin + SH2OMIN) ? (smcmin + SH2OMIN) : stotmin;
            if (stot < stotmin) {
                st->smc[k] = stotmin;
                wplus = (stot - stotmin) * ddz;
            } else {
                st->smc[k] = stot;
                wplus = 0.0;
            }
        }
        sh2omid[k] = st->smc[k] - sice[k];
    }
    ddz = -zsoil[0];
    for (k = 0; k < nsoil; k++) {
        if (k != 0) ddz = zsoil[k - 1] - zsoil[k];
        st->sh2o[k] = sh2omid[k] + wplus / ddz;
        stot = st->sh2o[k] + sice[k];
        wplus = (stot > smcmax) ? (stot - smcmax) * ddz : 0.0;
        st->smc[k] = (stot < smcmax) ? stot : smcmax;
        st->smc[k] = (st->smc[k] > smcmin + SH2OMIN) ? st->smc[k] : (smcmin + SH2OMIN);
        st->sh2o[k] = st->smc[k] - sice[k];
        st->sh2o[k] = (st->sh2o[k] > 0.0) ? st->sh2o[k] : 0.0;
    }
    st->runoff3 = wplus / dt;
}

/* SStep: src/noah/noah.c:2427-2481 */
static void SStep(const oracle_noah_col *in, const double *zsoil, oracle_noah_state *st, double *rhstt,
                  const double *sice, double *ai, double *bi, double *ci, double dt)
{
    int k, nsoil = in->nsoil;
    double rhsttin[NS_MAXLYR], ciin[NS_MAXLYR], sh2o0[NS_MAXLYR];
    for (k = 0; k < nsoil; k++) {
        rhstt[k] *= dt;
        ai[k] *= dt;
        bi[k] = 1.0 + bi[k] * dt;
        ci[k] *= dt;
        sh2o0[k] = st->sh2o[k];
        st->smflxv[k] = 0.0;
    }
    for (k = 0; k < nsoil; k++) {
        rhsttin[k] = rhstt[k];
        ciin[k] = ci[k];
    }
    oracle_Rosr12(ci, ai, bi, ciin, rhsttin, rhstt, nsoil);
    for (k = 0; k < nsoil; k++) st->sh2o[k] += ci[k];
    AdjSmProf(in, zsoil, sice, dt, st);
    for (k = nsoil - 1; k > 0; k--) {
        st->smflxv[k - 1] = (st->sh2o[k] - sh2o0[k]) * in->sldpth[k] / dt + st->runoff2_lyr[k] + in->et[k] +
            st->smflxv[k];
    }
}

void oracle_noah_soil_column(const oracle_noah_col *in, oracle_noah_state *st, double dt)
{
    int k, nsoil = in->nsoil;
    double zsoil[NS_MAXLYR], sice[NS_MAXLYR], ai[NS_MAXLYR], bi[NS_MAXLYR], ci[NS_MAXLYR], rhstt[NS_MAXLYR];
    /* layer depths below the surface, as the reference derives them from sldpth (src/noah/lsm_func.c:150-160) */
    zsoil[0] = -in->sldpth[0];
    for (k = 1; k < nsoil; k++) zsoil[k] = -in->sldpth[k] + zsoil[k - 1];
    /* Summary, src/update.c:37-43 */
    st->nwtbl = oracle_FindWaterTable(in->sldpth, nsoil, in->gw, st->satdpth);
    oracle_CalcLatFlx(nsoil, st->satdpth, in->runoff2, st->runoff2_lyr);
    /* NoahHydrol, src/noah/noah.c:57-94 */
    st->nwtbl = oracle_FindWaterTable(in->sldpth, nsoil, in->gw, st->satdpth);
    for (k = 0; k < nsoil; k++) {
        st->smc[k] = (st->smc[k] > in->smcmin + SH2OMIN) ? st->smc[k] : in->smcmin + SH2OMIN;
        st->smc[k] = (st->smc[k] < in->smcmax) ? st->smc[k] : in->smcmax;
        st->sh2o[k] = (st->sh2o[k] < st->smc[k]) ? st->sh2o[k] : st->smc[k];
    }
    /* SmFlx, src/noah/noah.c:1654-1712 */
    for (k = 0; k < nsoil; k++) sice[k] = st->smc[k] - st->sh2o[k];
    if (0 == st->nwtbl) {
        for (k = 0; k < nsoil; k++) {
            st->smc[k] = in->smcmax;
            st->sh2o[k] = st->smc[k] - sice[k];
        }
    } else {
        SRT(in, zsoil, st, rhstt, sice, ai, bi, ci);
        SStep(in, zsoil, st, rhstt, sice, ai, bi, ci, dt);
    }
}

void oracle_noah_soil_step(int ne, double dt, const int *nsoil, const double *sldpth, const double *smcmax,
                           const double *smcmin, const double *ksatv, const double *alpha, const double *beta,
                           const double *frzx, const double *gw, const double *infil, const double *runoff2,
                           const double *edir, const double *et, double *smc, double *sh2o, double *fcr,
                           int *nwtbl, double *satdpth, double *runoff2_lyr, double *smflxv, double *runoff3)
{
    int i;
#if defined(_OPENMP)
# pragma omp parallel for
#endif
    for (i = 0; i < ne; i++) {
        const size_t o = (size_t)i * NS_MAXLYR;
        oracle_noah_col in;
        oracle_noah_state st;
        in.nsoil = nsoil[i];
        in.sldpth = sldpth + o;
        in.smcmax = smcmax[i]; in.smcmin = smcmin[i]; in.ksatv = ksatv[i];
        in.alpha = alpha[i]; in.beta = beta[i]; in.frzx = frzx[i];
        in.gw = gw[i]; in.infil = infil[i]; in.runoff2 = runoff2[i]; in.edir = edir[i];
        in.et = et + o;
        memset(&st, 0, sizeof(st));
        memcpy(st.smc, smc + o, sizeof(st.smc));
        memcpy(st.sh2o, sh2o + o, sizeof(st.sh2o));
        memcpy(st.smflxv, smflxv + o, sizeof(st.smflxv));
        memcpy(st.runoff2_lyr, runoff2_lyr + o, sizeof(st.runoff2_lyr));
        st.fcr = fcr[i];
        st.runoff3 = runoff3[i];
        oracle_noah_soil_column(&in, &st, dt);
        memcpy(smc + o, st.smc, sizeof(st.smc));
        memcpy(sh2o + o, st.sh2o, sizeof(st.sh2o));
        memcpy(smflxv + o, st.smflxv, sizeof(st.smflxv));
        memcpy(satdpth + o, st.satdpth, sizeof(st.satdpth));
        memcpy(runoff2_lyr + o, st.runoff2_lyr, sizeof(st.runoff2_lyr));
        fcr[i] = st.fcr;
        runoff3[i] = st.runoff3;
        nwtbl[i] = st.nwtbl;
    }
}
