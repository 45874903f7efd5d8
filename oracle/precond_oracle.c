/*
 * precond_oracle.c - block-Jacobi preconditioner of the host model-step
 * backend: the host leg of the device preconditioner's parity tests.
 *
 * TEST INFRASTRUCTURE ONLY (tests/ and bench.py's CPU baseline).
 *
 * P = I - gamma * J_bd, J_bd the diagonal blocks of the RHS Jacobian: per
 * element the 3x3 block of (surf, unsat, gw), per river segment the 2x2 block
 * of (stage, gw), identity for any further state entries.  J_bd comes from
 * difference quotients, one RHS call per (colour, local variable): nodes of a
 * colour are pairwise more than two links apart in the graph of elements
 * (edge neighbours), river segments (downstream link) and their bank
 * elements, and an RHS component reads only states within two links, so each
 * column is that of a single-node perturbation.  Every call starts from the
 * same carried river-edge flux (the state the RHS carries between calls is
 * saved before and restored after each call).  The reference integrates
 * without a preconditioner (src/ode.c:523); this is the SPGMR preconditioning
 * of cvode_spgmr.c with its setup/solve callback shape (cvode_spils.h).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pihm_oracle.h"
#include "../include/pihm_cvode.h"

typedef struct oracle_bj {
    const pihm_mesh_desc *m;
    oracle_state *s;
    int ne, nr;
    long neq;
    int ncolor;
    int32_t *color;            /* [ne + nr], elements then rivers */
    double *J;                 /* [9 ne + 4 nr] */
    double *P;                 /* inverses of I - gamma J, same layout */
    double *yp, *f0, *fp, *ovl_save;
    long nsetup_j;             /* Jacobian evaluations */
} oracle_bj;

/* adjacency of the node graph, CSR */
static void build_graph(const pihm_mesh_desc *m, int **off_out, int **adj_out)
{
    const int ne = m->nelem, nr = m->nriver, n = ne + nr;
    int *deg = (int *)calloc((size_t)n, sizeof(int));
    int *off = (int *)calloc((size_t)n + 1, sizeof(int));
    int i, j, r, pass;
    int *adj = NULL, *fill = NULL;
    for (pass = 0; pass < 2; pass++) {
#define LINK(a, b) do { if (pass == 0) { deg[a]++; deg[b]++; } \
                         else { adj[fill[a]++] = (b); adj[fill[b]++] = (a); } } while (0)
        for (i = 0; i < ne; i++)
            for (j = 0; j < 3; j++) {
                const int c = m->nabr[3 * i + j];
                if (c > 0 && c - 1 > i) LINK(i, c - 1);
                else if (c < 0) LINK(i, ne + (-c - 1));
            }
        for (r = 0; r < nr; r++) {
            const int L = m->leftele[r] - 1, R = m->rightele[r] - 1, d = m->down[r];
            if (L >= 0) LINK(ne + r, L);
            if (R >= 0) LINK(ne + r, R);
            if (L >= 0 && R >= 0) LINK(L, R);
            if (d > 0) LINK(ne + r, ne + d - 1);
        }
#undef LINK
        if (pass == 0) {
            for (i = 0; i < n; i++) off[i + 1] = off[i] + deg[i];
            adj = (int *)malloc((size_t)off[n] * sizeof(int) + 1);
            fill = (int *)malloc((size_t)n * sizeof(int) + 1);
            memcpy(fill, off, (size_t)n * sizeof(int));
        }
    }
    free(deg);
    free(fill);
    *off_out = off;
    *adj_out = adj;
}

/* greedy distance-2 colouring in node order */
static int colour(const pihm_mesh_desc *m, int32_t *color)
{
    const int n = m->nelem + m->nriver;
    int *off, *adj, u, a, b, ncol = 0;
    int *mark;
    build_graph(m, &off, &adj);
    mark = (int *)malloc(((size_t)n + 1) * sizeof(int));
    for (u = 0; u <= n; u++) mark[u] = -1;
    for (u = 0; u < n; u++) color[u] = -1;
    for (u = 0; u < n; u++) {
        int c = 0;
        for (a = off[u]; a < off[u + 1]; a++) {
            const int v = adj[a];
            if (color[v] >= 0) mark[color[v]] = u;
            for (b = off[v]; b < off[v + 1]; b++) {
                const int w = adj[b];
                if (w != u && color[w] >= 0) mark[color[w]] = u;
            }
        }
        while (mark[c] == u) c++;
        color[u] = c;
        if (c + 1 > ncol) ncol = c + 1;
    }
    free(mark);
    free(off);
    free(adj);
    return ncol;
}

oracle_bj *oracle_bj_new(const pihm_mesh_desc *m, oracle_state *s, long neq)
{
    oracle_bj *p = (oracle_bj *)calloc(1, sizeof(*p));
    const int ne = m->nelem, nr = m->nriver;
    p->m = m; p->s = s; p->ne = ne; p->nr = nr; p->neq = neq;
    p->color = (int32_t *)malloc(((size_t)ne + nr + 1) * sizeof(int32_t));
    p->ncolor = colour(m, p->color);
    p->J = (double *)calloc(9 * (size_t)ne + 4 * (size_t)nr + 1, sizeof(double));
    p->P = (double *)calloc(9 * (size_t)ne + 4 * (size_t)nr + 1, sizeof(double));
    p->yp = (double *)malloc((size_t)neq * sizeof(double));
    p->f0 = (double *)malloc((size_t)neq * sizeof(double));
    p->fp = (double *)malloc((size_t)neq * sizeof(double));
    p->ovl_save = (double *)malloc(3 * (size_t)ne * sizeof(double) + 8);
    return p;
}

void oracle_bj_free(oracle_bj *p)
{
    if (!p) return;
    free(p->color); free(p->J); free(p->P); free(p->yp); free(p->f0); free(p->fp); free(p->ovl_save);
    free(p);
}

int oracle_bj_ncolor(const oracle_bj *p) { return p->ncolor; }
const int32_t *oracle_bj_colors(const oracle_bj *p) { return p->color; }
const double *oracle_bj_jac(const oracle_bj *p) { return p->J; }

/* state index of node u's local variable v (-1 past a river's two) */
static long state_index(const oracle_bj *p, int u, int v)
{
    if (u < p->ne) return (long)v * p->ne + u;
    if (v > 1) return -1;
    return 3L * p->ne + (long)v * p->nr + (u - p->ne);
}

static double dq_increment(double y)
{
    const double a = fabs(y);
    return sqrt(DBL_EPSILON) * (a > 1e-2 ? a : 1e-2);
}

/* J_bd by coloured difference quotients at (t, y); the carried river-edge
 * fluxes are left as they were.  Returns the RHS's code. */
int oracle_bj_jacobian(oracle_bj *p, double t, const double *y)
{
    const int ne = p->ne, nr = p->nr, n = ne + nr;
    const size_t novl = 3 * (size_t)ne * sizeof(double);
    int c, v, u, w, rv;
    memcpy(p->ovl_save, p->s->ovlflow, novl);
    rv = oracle_ode(p->m, p->s, t, y, p->f0, NULL);
    memcpy(p->s->ovlflow, p->ovl_save, novl);
    if (rv) return rv;
    for (c = 0; c < p->ncolor; c++)
        for (v = 0; v < 3; v++) {
            int any = 0;
            memcpy(p->yp, y, (size_t)p->neq * sizeof(double));
            for (u = 0; u < n; u++) {
                const long k = state_index(p, u, v);
                if (p->color[u] != c || k < 0) continue;
                p->yp[k] = y[k] + dq_increment(y[k]);
                any = 1;
            }
            if (!any) continue;
            rv = oracle_ode(p->m, p->s, t, p->yp, p->fp, NULL);
            memcpy(p->s->ovlflow, p->ovl_save, novl);
            if (rv) return rv;
            for (u = 0; u < n; u++) {
                const long k = state_index(p, u, v);
                const int nv = (u < ne) ? 3 : 2;
                double *Jb;
                if (p->color[u] != c || k < 0) continue;
                const double inc = p->yp[k] - y[k];
                Jb = (u < ne) ? p->J + 9L * u : p->J + 9L * ne + 4L * (u - ne);
                for (w = 0; w < nv; w++) {
                    const long q = state_index(p, u, w);
                    Jb[w * nv + v] = (p->fp[q] - p->f0[q]) / inc;
                }
            }
        }
    p->nsetup_j++;
    return 0;
}

/* P = inverse of (I - gamma J) per block; 1 when a block is singular */
int oracle_bj_factor(oracle_bj *p, double gamma)
{
    int i;
    for (i = 0; i < p->ne; i++) {
        const double *J = p->J + 9L * i;
        double *P = p->P + 9L * i;
        double a[9], det;
        int k;
        for (k = 0; k < 9; k++) a[k] = -gamma * J[k];
        a[0] += 1.0; a[4] += 1.0; a[8] += 1.0;
        P[0] = a[4] * a[8] - a[5] * a[7];
        P[1] = a[2] * a[7] - a[1] * a[8];
        P[2] = a[1] * a[5] - a[2] * a[4];
        P[3] = a[5] * a[6] - a[3] * a[8];
        P[4] = a[0] * a[8] - a[2] * a[6];
        P[5] = a[2] * a[3] - a[0] * a[5];
        P[6] = a[3] * a[7] - a[4] * a[6];
        P[7] = a[1] * a[6] - a[0] * a[7];
        P[8] = a[0] * a[4] - a[1] * a[3];
        det = a[0] * P[0] + a[1] * P[3] + a[2] * P[6];
        if (!(fabs(det) > 0.0) || !isfinite(det)) return 1;
        for (k = 0; k < 9; k++) P[k] /= det;
    }
    for (i = 0; i < p->nr; i++) {
        const double *J = p->J + 9L * p->ne + 4L * i;
        double *P = p->P + 9L * p->ne + 4L * i;
        const double a0 = 1.0 - gamma * J[0], a1 = -gamma * J[1], a2 = -gamma * J[2], a3 = 1.0 - gamma * J[3];
        const double det = a0 * a3 - a1 * a2;
        if (!(fabs(det) > 0.0) || !isfinite(det)) return 1;
        P[0] = a3 / det; P[1] = -a1 / det; P[2] = -a2 / det; P[3] = a0 / det;
    }
    return 0;
}

/* z = P r (block-wise; identity past the rivers) */
void oracle_bj_apply(const oracle_bj *p, const double *r, double *z)
{
    const int ne = p->ne, nr = p->nr;
    long k;
    int i;
    for (i = 0; i < ne; i++) {
        const double *P = p->P + 9L * i;
        const double r0 = r[i], r1 = r[ne + i], r2 = r[2L * ne + i];
        z[i] = P[0] * r0 + P[1] * r1 + P[2] * r2;
        z[ne + i] = P[3] * r0 + P[4] * r1 + P[5] * r2;
        z[2L * ne + i] = P[6] * r0 + P[7] * r1 + P[8] * r2;
    }
    for (i = 0; i < nr; i++) {
        const double *P = p->P + 9L * ne + 4L * i;
        const long a = 3L * ne + i, b = 3L * ne + nr + i;
        const double r0 = r[a], r1 = r[b];
        z[a] = P[0] * r0 + P[1] * r1;
        z[b] = P[2] * r0 + P[3] * r1;
    }
    for (k = 3L * ne + 2L * nr; k < p->neq; k++) z[k] = r[k];
}

/* PihmPrecSetupFn / PihmPrecSolveFn over host vectors */
int oracle_bj_psetup(realtype t, N_Vector y, N_Vector fy, int jok, int *jcur, realtype gamma, void *ud,
                     N_Vector t1, N_Vector t2, N_Vector t3)
{
    oracle_bj *p = (oracle_bj *)ud;
    (void)fy; (void)t1; (void)t2; (void)t3;
    if (!jok) {
        const int rv = oracle_bj_jacobian(p, t, y->ops->nvgetarraypointer(y));
        if (rv) return rv < 0 ? -1 : 1;
        *jcur = 1;
    } else {
        *jcur = 0;
    }
    return oracle_bj_factor(p, gamma);
}

int oracle_bj_psolve(realtype t, N_Vector y, N_Vector fy, N_Vector r, N_Vector z, realtype gamma, realtype delta,
                     int lr, void *ud, N_Vector tmp)
{
    (void)t; (void)y; (void)fy; (void)gamma; (void)delta; (void)lr; (void)tmp;
    oracle_bj_apply((const oracle_bj *)ud, r->ops->nvgetarraypointer(r), z->ops->nvgetarraypointer(z));
    return 0;
}
