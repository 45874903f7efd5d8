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
#include "../include/pihm_driver.h"

N_Vector oracle_nv_new(long int length, int nthreads);
void oracle_nv_destroy(N_Vector v);
double *oracle_nv_data(N_Vector v);

typedef struct oracle_backend {
    const pihm_mesh_desc *m;
    oracle_state *s;
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

static int oracle_summary(void *user, N_Vector y)
{
    oracle_backend *b = (oracle_backend *)user;
    memcpy(b->s->ws0_surf, y->ops->nvgetarraypointer(y), b->m->nelem * sizeof(double));
    return 0;
}

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

const pihm_backend *oracle_backend_get(oracle_backend *b) { return &b->be; }

void oracle_backend_free(oracle_backend *b)
{
    if (!b) return;
    free(b->prcp); free(b->edir); free(b->ett); free(b->station);
    free(b);
}

/* y as a host vector for the driver; data[] in caller order */
N_Vector oracle_backend_vector(long n, int nthreads) { return oracle_nv_new(n, nthreads); }
