/*
 * oracle_backend.c - host backend of the model-step driver built on the
 * oracle RHS and the host N_Vector.
 *
 * TEST INFRASTRUCTURE ONLY: the reference-side (CPU) leg of the integration
 * parity tests and the timed CPU baseline.  Forcing interpolation is the
 * same arithmetic as the device backend (IntrplForc, src/forcing.c:514-519).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pihm_oracle.h"
#include "noahlsm_oracle.h"
#include "../include/pihm_driver.h"

N_Vector oracle_nv_new(long int length, int nthreads);
void oracle_nv_destroy(N_Vector v);
double *oracle_nv_data(N_Vector v);

typedef struct oracle_bj oracle_bj;
oracle_bj *oracle_bj_new(const pihm_mesh_desc *m, oracle_state *s, long neq);
void oracle_bj_free(oracle_bj *p);
int oracle_bj_psetup(realtype t, N_Vector y, N_Vector fy, int jok, int *jcur, realtype gamma, void *ud,
                     N_Vector t1, N_Vector t2, N_Vector t3);
int oracle_bj_psolve(realtype t, N_Vector y, N_Vector fy, N_Vector r, N_Vector z, realtype gamma, realtype delta,
                     int lr, void *ud, N_Vector tmp);

typedef struct oracle_noah_soil_host {
    int *nsoil, *nroot;
    double *sldpth, *smcmax, *frzx, *et;              /* per-layer arrays [i*11 + k] */
    double *smc, *sh2o, *smflxv, *runoff2_lyr, *runoff3;
    double *prev_u, *prev_g, *ydot, *ovl_save;
    int started;
} oracle_noah_soil_host;

/* The Noah land-surface step at every LSM step (caller's arrays, which must
 * outlive the backend): parameters and carried state per element, station
 * records (7 variables, record-major), the element's station, projected LAI. */
typedef struct oracle_noah_lsm_host {
    const oracle_noahlsm_par *par;
    oracle_noahlsm_state *st;
    const int *station;
    int nstation, nrec;
    const double *ftime, *data, *lai;
    double etstep;
    double *met;
    oracle_noahlsm_out *out;
} oracle_noah_lsm_host;

typedef struct oracle_backend {
    const pihm_mesh_desc *m;
    oracle_state *s;
    oracle_bj *bj;
    oracle_noah_soil_host *ns;
    oracle_noah_lsm_host *nl;
    int nstation, nrec;
    double rec_dt;
    double *prcp, *edir, *ett;
    int32_t *station;
    int nthreads;
    pihm_backend be;
} oracle_backend;

/* CVRhsFn over the oracle (src/ode.c:4 signature) */
static int oracle_cvrhs(double t, N_Vector y, N_Vector ydot, void *ud)
{
    oracle_backend *b = (oracle_backend *)ud;
    double *yd = y->ops->nvgetarraypointer(y);
    double *dd = ydot->ops->nvgetarraypointer(ydot);
    return oracle_ode(b->m, b->s, t, yd, dd, NULL);
}

static int oracle_forcing(void *user, double t)
{
    oracle_backend *b = (oracle_backend *)user;
    long k = (long)floor(t / b->rec_dt);
    int i;
    if (k < 0) k = 0;
    if (k > b->nrec - 2) k = b->nrec - 2;
    double ta = (double)k * b->rec_dt, tb = (double)(k + 1) * b->rec_dt;
    double ed = ((tb - t) * b->edir[k] + (t - ta) * b->edir[k + 1]) / (tb - ta);
    double et = ((tb - t) * b->ett[k] + (t - ta) * b->ett[k + 1]) / (tb - ta);
#if defined(_OPENMP)
# pragma omp parallel for num_threads(b->nthreads)
#endif
    for (i = 0; i < b->m->nelem; i++) {
        const double *P = b->prcp + (long)b->station[i] * b->nrec;
        b->s->pcpdrp[i] = ((tb - t) * P[k] + (t - ta) * P[k + 1]) / (tb - ta);
        b->s->edir[i] = ed;
        b->s->ett[i] = et;
    }
    return 0;
}

/* Summary with _NOAH_ (src/update.c:3-60) and NoahHydrol (src/pihm.c:106-109):
 * the fluxes at y (the carried river-edge flux kept), MassBalance with the
 * _NOAH_ EtExtract split, the soil column, and what the next RHS reads. */
static int oracle_noah_summary(oracle_backend *b, const double *y)
{
    oracle_noah_soil_host *n = b->ns;
    oracle_state *s = b->s;
    const pihm_mesh_desc *m = b->m;
    const int ne = m->nelem;
    int i;
    if (!n->started) {
        for (i = 0; i < ne; i++) { n->prev_u[i] = y[ne + i]; n->prev_g[i] = y[2 * ne + i]; }
        n->started = 1;
        return 0;
    }
    memcpy(n->ovl_save, s->ovlflow, 3 * (size_t)ne * sizeof(double));
    oracle_ode(m, s, 0.0, y, n->ydot, NULL);
    memcpy(s->ovlflow, n->ovl_save, 3 * (size_t)ne * sizeof(double));
    for (i = 0; i < ne; i++) {
        const size_t o = (size_t)i * 11;
        double soilw0 = n->prev_g[i] + n->prev_u[i], soilw1 = y[2 * ne + i] + y[ne + i], subrunoff = 0.0, infil;
        int j;
        oracle_noah_col in;
        oracle_noah_state st;
        soilw0 = (soilw0 > m->depth[i]) ? m->depth[i] : soilw0;
        soilw0 = (soilw0 < 0.0) ? 0.0 : soilw0;
        soilw1 = (soilw1 > m->depth[i]) ? m->depth[i] : soilw1;
        soilw1 = (soilw1 < 0.0) ? 0.0 : soilw1;
        for (j = 0; j < 3; j++) subrunoff += s->subsurf[(size_t)i * 3 + j] / m->area[i];
        infil = (soilw1 - soilw0) * m->porosity[i] / m->stepsize + subrunoff + s->edir_unsat[i] + s->edir_gw[i] +
            s->ett_unsat[i] + s->ett_gw[i];
        if (infil < 0.0) { subrunoff -= infil; infil = 0.0; }
        in.nsoil = n->nsoil[i];
        in.sldpth = n->sldpth + o;
        in.smcmax = n->smcmax[i]; in.smcmin = s->smcmin[i]; in.ksatv = m->ksatv[i];
        in.alpha = m->alpha[i]; in.beta = m->beta[i]; in.frzx = n->frzx[i];
        in.gw = y[2 * ne + i]; in.infil = infil; in.runoff2 = subrunoff; in.edir = s->edir[i];
        in.et = n->et + o;
        memset(&st, 0, sizeof(st));
        memcpy(st.smc, n->smc + o, sizeof(st.smc));
        memcpy(st.sh2o, n->sh2o + o, sizeof(st.sh2o));
        memcpy(st.smflxv, n->smflxv + o, sizeof(st.smflxv));
        memcpy(st.runoff2_lyr, n->runoff2_lyr + o, sizeof(st.runoff2_lyr));
        st.fcr = s->fcr[i];
        st.runoff3 = n->runoff3[i];
        oracle_noah_soil_column(&in, &st, m->stepsize);
        memcpy(n->smc + o, st.smc, sizeof(st.smc));
        memcpy(n->sh2o + o, st.sh2o, sizeof(st.sh2o));
        memcpy(n->smflxv + o, st.smflxv, sizeof(st.smflxv));
        memcpy(n->runoff2_lyr + o, st.runoff2_lyr, sizeof(st.runoff2_lyr));
        n->runoff3[i] = st.runoff3;
        s->sh2o0[i] = st.sh2o[0];
        s->fcr[i] = st.fcr;
        s->nwtbl[i] = st.nwtbl;
        n->prev_u[i] = y[ne + i];
        n->prev_g[i] = y[2 * ne + i];
    }
    return 0;
}

static int oracle_summary(void *user, N_Vector y)
{
    oracle_backend *b = (oracle_backend *)user;
    if (b->ns) {
        const int rc = oracle_noah_summary(b, y->ops->nvgetarraypointer(y));
        if (rc) return rc;
    }
    memcpy(b->s->ws0_surf, y->ops->nvgetarraypointer(y), b->m->nelem * sizeof(double));
    return 0;
}

/* The soil-water step at every Summary, over the state's _NOAH_ RHS branches
 * (oracle_set_noah with 11-layer et); per-layer arrays [i*11 + k]. */
void oracle_backend_set_noah_soil(oracle_backend *b, const int *nsoil, const double *sldpth, const double *smcmax,
                                  const double *frzx, const double *smc, const double *sh2o)
{
    const size_t ne = (size_t)b->m->nelem, nl = ne * 11;
    oracle_noah_soil_host *n = (oracle_noah_soil_host *)calloc(1, sizeof(*n));
    n->nsoil = (int *)malloc(ne * sizeof(int));
    memcpy(n->nsoil, nsoil, ne * sizeof(int));
    n->sldpth = (double *)malloc(nl * sizeof(double)); memcpy(n->sldpth, sldpth, nl * sizeof(double));
    n->smcmax = (double *)malloc(ne * sizeof(double)); memcpy(n->smcmax, smcmax, ne * sizeof(double));
    n->frzx = (double *)malloc(ne * sizeof(double)); memcpy(n->frzx, frzx, ne * sizeof(double));
    n->smc = (double *)malloc(nl * sizeof(double)); memcpy(n->smc, smc, nl * sizeof(double));
    n->sh2o = (double *)malloc(nl * sizeof(double)); memcpy(n->sh2o, sh2o, nl * sizeof(double));
    n->et = b->s->et;     /* the RHS's per-layer transpiration, [i*11 + k] */
    n->smflxv = (double *)calloc(nl, sizeof(double));
    n->runoff2_lyr = (double *)calloc(nl, sizeof(double));
    n->runoff3 = (double *)calloc(ne, sizeof(double));
    n->prev_u = (double *)calloc(ne, sizeof(double));
    n->prev_g = (double *)calloc(ne, sizeof(double));
    n->ydot = (double *)calloc((size_t)b->m->nelem * 3 + 2 * (size_t)b->m->nriver + 8, sizeof(double));
    n->ovl_save = (double *)calloc(3 * ne, sizeof(double));
    b->ns = n;
}

/* smc and sh2o [i*11 + k] of the soil-water step */
void oracle_backend_noah_soil_state(const oracle_backend *b, double *smc, double *sh2o)
{
    const size_t nl = (size_t)b->m->nelem * 11;
    if (!b->ns) return;
    memcpy(smc, b->ns->smc, nl * sizeof(double));
    memcpy(sh2o, b->ns->sh2o, nl * sizeof(double));
}

int oracle_meteo_interp_n(const double *ftime, const double *data, int nstation, int nrec, int nvar, double t,
                          double *met);   /* lsm_oracle.c */

/* Noah() every LSM step (src/pihm.c:25-63 with _NOAH_) on the soil columns of
 * the soil-water step: the stations' records at t, the element's forcing as
 * ApplyMeteoForc sets it (src/forcing.c:340-353), then what the RHS reads. */
static int oracle_noah_forcing(void *user, double t)
{
    oracle_backend *b = (oracle_backend *)user;
    oracle_noah_lsm_host *n = b->nl;
    oracle_noah_soil_host *ns = b->ns;
    oracle_state *s = b->s;
    const int ne = b->m->nelem;
    int i;
    if (oracle_meteo_interp_n(n->ftime, n->data, n->nstation, n->nrec, 7, t, n->met)) return -1;
#if defined(_OPENMP)
# pragma omp parallel for num_threads(b->nthreads)
#endif
    for (i = 0; i < ne; i++) {
        const size_t o = (size_t)i * 11;
        const int k = n->station[i];
        oracle_noahlsm_forc f;
        int j;
        f.prcp = n->met[0 * n->nstation + k] / 1000.0;
        f.sfctmp = n->met[1 * n->nstation + k];
        f.rh = n->met[2 * n->nstation + k];
        f.sfcspd = n->met[3 * n->nstation + k];
        f.soldn = n->met[4 * n->nstation + k];
        f.soldn = (f.soldn > 0.0) ? f.soldn : 0.0;
        f.longwave = n->met[5 * n->nstation + k];
        f.sfcprs = n->met[6 * n->nstation + k];
        f.proj_lai = n->lai[i];
        memcpy(n->st[i].smc, ns->smc + o, sizeof(n->st[i].smc));
        memcpy(n->st[i].sh2o, ns->sh2o + o, sizeof(n->st[i].sh2o));
        oracle_noah_element(&n->par[i], &n->st[i], &f, n->etstep, &n->out[i]);
        memcpy(ns->smc + o, n->st[i].smc, sizeof(n->st[i].smc));
        memcpy(ns->sh2o + o, n->st[i].sh2o, sizeof(n->st[i].sh2o));
        s->pcpdrp[i] = n->out[i].pcpdrp;
        s->edir[i] = n->out[i].edir;
        s->ett[i] = n->out[i].ett;
        for (j = 0; j < 11; j++) s->et[o + j] = n->out[i].et[j];
        s->sh2o0[i] = n->st[i].sh2o[0];
    }
    return 0;
}

void oracle_backend_set_noah_lsm(oracle_backend *b, const oracle_noahlsm_par *par, oracle_noahlsm_state *st,
                                 const int *station, int nstation, int nrec, const double *ftime,
                                 const double *data, const double *lai, double etstep)
{
    oracle_noah_lsm_host *n = (oracle_noah_lsm_host *)calloc(1, sizeof(*n));
    n->par = par; n->st = st; n->station = station; n->nstation = nstation; n->nrec = nrec;
    n->ftime = ftime; n->data = data; n->lai = lai; n->etstep = etstep;
    n->met = (double *)calloc((size_t)7 * nstation, sizeof(double));
    n->out = (oracle_noahlsm_out *)calloc((size_t)b->m->nelem, sizeof(oracle_noahlsm_out));
    b->nl = n;
    b->be.apply_forcing = oracle_noah_forcing;
}

/* the initial Summary (the storage the first balance starts from) */
int oracle_backend_start(oracle_backend *b, N_Vector y) { return oracle_summary(b, y); }

oracle_backend *oracle_backend_new(const pihm_mesh_desc *m, oracle_state *s,
                                   const pihm_forcing *f, int nthreads)
{
    oracle_backend *b = (oracle_backend *)calloc(1, sizeof(*b));
    size_t np = (size_t)f->nstation * f->nrec;
    b->m = m; b->s = s;
    b->nstation = f->nstation; b->nrec = f->nrec; b->rec_dt = f->rec_dt;
    b->prcp = (double *)malloc(np * sizeof(double));
    b->edir = (double *)malloc(f->nrec * sizeof(double));
    b->ett = (double *)malloc(f->nrec * sizeof(double));
    b->station = (int32_t *)malloc(m->nelem * sizeof(int32_t));
    memcpy(b->prcp, f->prcp, np * sizeof(double));
    memcpy(b->edir, f->edir, f->nrec * sizeof(double));
    memcpy(b->ett, f->ett, f->nrec * sizeof(double));
    memcpy(b->station, f->station, m->nelem * sizeof(int32_t));
    b->nthreads = nthreads > 0 ? nthreads : 1;
    b->be.rhs = oracle_cvrhs;
    b->be.rhs_data = b;
    b->be.apply_forcing = oracle_forcing;
    b->be.summary = oracle_summary;
    b->be.model_step_end = NULL;
    b->be.user = b;
    return b;
}

/* the RHS's user data is the backend: hand the preconditioner its own data */
static int backend_psetup(realtype t, N_Vector y, N_Vector fy, int jok, int *jcur, realtype gamma, void *ud,
                          N_Vector t1, N_Vector t2, N_Vector t3)
{
    return oracle_bj_psetup(t, y, fy, jok, jcur, gamma, ((oracle_backend *)ud)->bj, t1, t2, t3);
}

static int backend_psolve(realtype t, N_Vector y, N_Vector fy, N_Vector r, N_Vector z, realtype gamma,
                          realtype delta, int lr, void *ud, N_Vector tmp)
{
    return oracle_bj_psolve(t, y, fy, r, z, gamma, delta, lr, ((oracle_backend *)ud)->bj, tmp);
}

/* Block-Jacobi preconditioner for runs with cfg->precond (neq: the state length) */
void oracle_backend_enable_precond(oracle_backend *b, long neq)
{
    if (!b->bj) b->bj = oracle_bj_new(b->m, b->s, neq);
    b->be.prec_setup = backend_psetup;
    b->be.prec_solve = backend_psolve;
}

oracle_bj *oracle_backend_bj(oracle_backend *b) { return b->bj; }

const pihm_backend *oracle_backend_get(oracle_backend *b) { return &b->be; }

void oracle_backend_free(oracle_backend *b)
{
    if (b && b->ns) {
        oracle_noah_soil_host *n = b->ns;
        free(n->nsoil); free(n->sldpth); free(n->smcmax); free(n->frzx); free(n->smc); free(n->sh2o);
        free(n->smflxv); free(n->runoff2_lyr); free(n->runoff3); free(n->prev_u); free(n->prev_g);
        free(n->ydot); free(n->ovl_save); free(n);
        b->ns = NULL;
    }
    if (!b) return;
    if (b->nl) {
        free(b->nl->met); free(b->nl->out); free(b->nl);
        b->nl = NULL;
    }
    oracle_bj_free(b->bj);
    free(b->prcp); free(b->edir); free(b->ett); free(b->station);
    free(b);
}

/* y as a host vector for the driver; data[] in caller order */
N_Vector oracle_backend_vector(long n, int nthreads) { return oracle_nv_new(n, nthreads); }

/* ------------------------------------------------------------------ land-surface backend
 * apply_forcing runs the land-surface step (oracle_meteo_interp, oracle_intcp_snow_et) on
 * the state the last summary saw, over etstep, as the reference's host path calls
 * IntcpSnowEt (src/pihm.c:30-63).  The descriptor, records, lai and z0 arrays are the
 * caller's and must outlive the backend. */
#include "../include/pihm_lsm.h"

int oracle_meteo_interp(const pihm_meteo *m, double t, double *met);
void oracle_intcp_snow_et(int ne, const pihm_lsm_desc *d, const double *met, int nstation,
                          const double *lai_e, const double *z0_e, double meltf,
                          const double *unsat_e, const double *gw_e, double stepsize,
                          double *cmc_e, double *sneqv_e, double *pcpdrp_e, double *edir_e,
                          double *ett_e, double *ec_e, double *drip_e);

typedef struct oracle_lsm_backend {
    const pihm_mesh_desc *m;
    oracle_state *s;
    const pihm_lsm_desc *d;
    const pihm_meteo *met;
    const double *lai, *z0;
    double meltf, etstep;
    double *cmc, *sneqv, *ec, *drip, *metv;
    const double *y;
    const pihm_ts *ets, *rts;              /* boundary series, optional */
    const int32_t *etype, *rtype;
    pihm_backend be;
} oracle_lsm_backend;

int oracle_bc_apply(const pihm_ts *ets, const int32_t *etype, int ne, double *elem_bc,
                    const pihm_ts *rts, const int32_t *rtype, int nr, double *river_bc, double t);

static int oracle_lsm_bc(void *user, double t)
{
    oracle_lsm_backend *b = (oracle_lsm_backend *)user;
    return oracle_bc_apply(b->ets, b->etype, b->m->nelem, b->s->elem_bc, b->rts, b->rtype, b->m->nriver,
                           b->s->river_bc, t) ? PIHM_ERR_ARG : 0;
}

static int oracle_cvrhs_lsm(double t, N_Vector y, N_Vector ydot, void *ud)
{
    oracle_lsm_backend *b = (oracle_lsm_backend *)ud;
    return oracle_ode(b->m, b->s, t, y->ops->nvgetarraypointer(y), ydot->ops->nvgetarraypointer(ydot), NULL);
}

static int oracle_lsm_forcing(void *user, double t)
{
    oracle_lsm_backend *b = (oracle_lsm_backend *)user;
    const int ne = b->m->nelem;
    if (!b->y || oracle_meteo_interp(b->met, t, b->metv)) return PIHM_ERR_ARG;
    oracle_intcp_snow_et(ne, b->d, b->metv, b->met->nstation, b->lai, b->z0, b->meltf, b->y + ne,
                         b->y + 2 * ne, b->etstep, b->cmc, b->sneqv, b->s->pcpdrp, b->s->edir,
                         b->s->ett, b->ec, b->drip);
    return 0;
}

static int oracle_lsm_summary(void *user, N_Vector y)
{
    oracle_lsm_backend *b = (oracle_lsm_backend *)user;
    b->y = y->ops->nvgetarraypointer(y);
    memcpy(b->s->ws0_surf, b->y, b->m->nelem * sizeof(double));
    return 0;
}

oracle_lsm_backend *oracle_lsm_backend_new(const pihm_mesh_desc *m, oracle_state *s,
                                           const pihm_lsm_desc *d, const pihm_meteo *met,
                                           const double *lai, const double *z0, double meltf,
                                           const double *cmc0, const double *sneqv0, double etstep)
{
    oracle_lsm_backend *b = (oracle_lsm_backend *)calloc(1, sizeof(*b));
    const size_t ne = (size_t)m->nelem;
    b->m = m; b->s = s; b->d = d; b->met = met;
    b->lai = lai; b->z0 = z0; b->meltf = meltf; b->etstep = etstep;
    b->cmc = (double *)malloc(ne * sizeof(double));
    b->sneqv = (double *)malloc(ne * sizeof(double));
    b->ec = (double *)calloc(ne, sizeof(double));
    b->drip = (double *)calloc(ne, sizeof(double));
    b->metv = (double *)calloc((size_t)met->nstation * PIHM_MET_NVAR, sizeof(double));
    memcpy(b->cmc, cmc0, ne * sizeof(double));
    memcpy(b->sneqv, sneqv0, ne * sizeof(double));
    b->be.rhs = oracle_cvrhs_lsm;
    b->be.rhs_data = b;
    b->be.apply_forcing = oracle_lsm_forcing;
    b->be.summary = oracle_lsm_summary;
    b->be.model_step_end = NULL;
    b->be.user = b;
    return b;
}

/* Boundary series applied every model step (ApplyBc, src/pihm.c:17-24). */
void oracle_lsm_backend_set_bc(oracle_lsm_backend *b, const pihm_ts *ets, const int32_t *etype,
                               const pihm_ts *rts, const int32_t *rtype)
{
    b->ets = ets; b->etype = etype; b->rts = rts; b->rtype = rtype;
    b->be.apply_bc = oracle_lsm_bc;
}

/* the initial Summary (ws0 = y0) before the driver runs */
int oracle_lsm_backend_start(oracle_lsm_backend *b, N_Vector y) { return oracle_lsm_summary(b, y); }

const pihm_backend *oracle_lsm_backend_get(oracle_lsm_backend *b) { return &b->be; }

void oracle_lsm_backend_state(const oracle_lsm_backend *b, double *cmc, double *sneqv)
{
    memcpy(cmc, b->cmc, b->m->nelem * sizeof(double));
    memcpy(sneqv, b->sneqv, b->m->nelem * sizeof(double));
}

void oracle_lsm_backend_free(oracle_lsm_backend *b)
{
    if (!b) return;
    free(b->cmc); free(b->sneqv); free(b->ec); free(b->drip); free(b->metv);
    free(b);
}
