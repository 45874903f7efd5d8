/*
 * nvector_host.c - host N_Vector restating SUNDIALS 2.7 nvector_serial.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the device N_Vector operations
 * and the vector type of the timed CPU baseline.
 *
 * Each operation follows cvode/src/nvec_ser/nvector_serial.c (line given at
 * each function), including the special cases of N_VLinearSum (:421-506)
 * whose rounding differs from the general formula (a == b, a == -b).
 * With nthreads > 1 the loops run under OpenMP like
 * cvode/src/nvec_openmp/nvector_openmp.c; the sums then depend on the thread
 * count, as there, but not on the order in which the threads finish.
 */
#include "../include/pihm_nvector.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#if defined(_OPENMP)
#include <omp.h>
#endif

typedef struct {
    long int length;
    int own_data;
    double *data;
    int nthreads;
} host_content;

#define HC(v) ((host_content *)((v)->content))
#define HD(v) (HC(v)->data)
#define HL(v) (HC(v)->length)
#define NT(v) (HC(v)->nthreads)

static N_Vector_ID host_id(N_Vector v) { (void)v; return SUNDIALS_NVEC_SERIAL; }
static N_Vector host_clone(N_Vector w);
static N_Vector host_cloneempty(N_Vector w);
void oracle_nv_destroy(N_Vector v);

static void host_space(N_Vector v, long int *lrw, long int *liw)
{ *lrw = HL(v); *liw = 1; }
static double *host_getarray(N_Vector v) { return HD(v); }
static void host_setarray(double *d, N_Vector v) { if (HL(v) > 0) HD(v) = d; }

/* N_VLinearSum_Serial :421-506 */
static void host_linearsum(double a, N_Vector x, double b, N_Vector y, N_Vector z)
{
    long int i, N = HL(x);
    double *xd = HD(x), *yd = HD(y), *zd = HD(z);
    int nt = NT(x);
    (void)nt;
    if ((b == 1.0) && (z == y)) {            /* Vaxpy(a, x, y) */
        if (a == 1.0) {
#pragma omp parallel for num_threads(nt) if (nt > 1)
            for (i = 0; i < N; i++) yd[i] += xd[i];
        } else if (a == -1.0) {
#pragma omp parallel for num_threads(nt) if (nt > 1)
            for (i = 0; i < N; i++) yd[i] -= xd[i];
        } else {
#pragma omp parallel for num_threads(nt) if (nt > 1)
            for (i = 0; i < N; i++) yd[i] += a * xd[i];
        }
        return;
    }
    if ((a == 1.0) && (z == x)) {            /* Vaxpy(b, y, x) */
        if (b == 1.0) {
#pragma omp parallel for num_threads(nt) if (nt > 1)
            for (i = 0; i < N; i++) xd[i] += yd[i];
        } else if (b == -1.0) {
#pragma omp parallel for num_threads(nt) if (nt > 1)
            for (i = 0; i < N; i++) xd[i] -= yd[i];
        } else {
#pragma omp parallel for num_threads(nt) if (nt > 1)
            for (i = 0; i < N; i++) xd[i] += b * yd[i];
        }
        return;
    }
    if (a == 1.0 && b == 1.0) {
#pragma omp parallel for num_threads(nt) if (nt > 1)
        for (i = 0; i < N; i++) zd[i] = xd[i] + yd[i];
        return;
    }
    if ((a == 1.0 && b == -1.0) || (a == -1.0 && b == 1.0)) {
        int test = (a == 1.0);
        double *v1 = test ? yd : xd, *v2 = test ? xd : yd;
#pragma omp parallel for num_threads(nt) if (nt > 1)
        for (i = 0; i < N; i++) zd[i] = v2[i] - v1[i];
        return;
    }
    if (a == 1.0 || b == 1.0) {
        int test = (a == 1.0);
        double c = test ? b : a;
        double *v1 = test ? yd : xd, *v2 = test ? xd : yd;
#pragma omp parallel for num_threads(nt) if (nt > 1)
        for (i = 0; i < N; i++) zd[i] = (c * v1[i]) + v2[i];
        return;
    }
    if (a == -1.0 || b == -1.0) {
        int test = (a == -1.0);
        double c = test ? b : a;
        double *v1 = test ? yd : xd, *v2 = test ? xd : yd;
#pragma omp parallel for num_threads(nt) if (nt > 1)
        for (i = 0; i < N; i++) zd[i] = (c * v1[i]) - v2[i];
        return;
    }
    if (a == b) {
#pragma omp parallel for num_threads(nt) if (nt > 1)
        for (i = 0; i < N; i++) zd[i] = a * (xd[i] + yd[i]);
        return;
    }
    if (a == -b) {
#pragma omp parallel for num_threads(nt) if (nt > 1)
        for (i = 0; i < N; i++) zd[i] = a * (xd[i] - yd[i]);
        return;
    }
#pragma omp parallel for num_threads(nt) if (nt > 1)
    for (i = 0; i < N; i++) zd[i] = (a * xd[i]) + (b * yd[i]);
}

/* N_VConst_Serial :508 */
static void host_const(double c, N_Vector z)
{
    long int i, N = HL(z); double *zd = HD(z); int nt = NT(z); (void)nt;
#pragma omp parallel for num_threads(nt) if (nt > 1)
    for (i = 0; i < N; i++) zd[i] = c;
}
/* N_VProd_Serial */
static void host_prod(N_Vector x, N_Vector y, N_Vector z)
{
    long int i, N = HL(x); double *xd = HD(x), *yd = HD(y), *zd = HD(z); int nt = NT(x); (void)nt;
#pragma omp parallel for num_threads(nt) if (nt > 1)
    for (i = 0; i < N; i++) zd[i] = xd[i] * yd[i];
}
/* N_VDiv_Serial */
static void host_div(N_Vector x, N_Vector y, N_Vector z)
{
    long int i, N = HL(x); double *xd = HD(x), *yd = HD(y), *zd = HD(z); int nt = NT(x); (void)nt;
#pragma omp parallel for num_threads(nt) if (nt > 1)
    for (i = 0; i < N; i++) zd[i] = xd[i] / yd[i];
}
/* N_VScale_Serial (z == x: scale in place; c == 1 copy; c == -1 negate) */
static void host_scale(double c, N_Vector x, N_Vector z)
{
    long int i, N = HL(x); double *xd = HD(x), *zd = HD(z); int nt = NT(x); (void)nt;
    if (z == x) {
#pragma omp parallel for num_threads(nt) if (nt > 1)
        for (i = 0; i < N; i++) xd[i] *= c;
        return;
    }
    if (c == 1.0) {
#pragma omp parallel for num_threads(nt) if (nt > 1)
        for (i = 0; i < N; i++) zd[i] = xd[i];
    } else if (c == -1.0) {
#pragma omp parallel for num_threads(nt) if (nt > 1)
        for (i = 0; i < N; i++) zd[i] = -xd[i];
    } else {
#pragma omp parallel for num_threads(nt) if (nt > 1)
        for (i = 0; i < N; i++) zd[i] = c * xd[i];
    }
}
static void host_abs(N_Vector x, N_Vector z)
{
    long int i, N = HL(x); double *xd = HD(x), *zd = HD(z); int nt = NT(x); (void)nt;
#pragma omp parallel for num_threads(nt) if (nt > 1)
    for (i = 0; i < N; i++) zd[i] = fabs(xd[i]);
}
static void host_inv(N_Vector x, N_Vector z)
{
    long int i, N = HL(x); double *xd = HD(x), *zd = HD(z); int nt = NT(x); (void)nt;
#pragma omp parallel for num_threads(nt) if (nt > 1)
    for (i = 0; i < N; i++) zd[i] = 1.0 / xd[i];
}
static void host_addconst(N_Vector x, double b, N_Vector z)
{
    long int i, N = HL(x); double *xd = HD(x), *zd = HD(z); int nt = NT(x); (void)nt;
#pragma omp parallel for num_threads(nt) if (nt > 1)
    for (i = 0; i < N; i++) zd[i] = xd[i] + b;
}
/* With nthreads > 1 a sum is split into one contiguous block per thread, each
 * summed left to right, and the blocks' sums are added in thread order: an OpenMP
 * reduction clause would add them in the order the threads finish, so two runs
 * on the same inputs could differ in the last bit and the integrator's step
 * choices after them.  Up to HOST_MAXT threads take part. */
#define HOST_MAXT 64
#define HOST_BLOCK_SUM(N, nt, TERM, out)                                          \
    do {                                                                           \
        double part_[HOST_MAXT];                                                   \
        int used_ = 1, t_;                                                         \
        const int want_ = (nt) < 1 ? 1 : ((nt) > HOST_MAXT ? HOST_MAXT : (nt));   \
        _Pragma("omp parallel num_threads(want_) if (want_ > 1)")                  \
        {                                                                          \
            int me_ = 0, all_ = 1;                                                 \
            long int i, lo_, hi_;                                                  \
            double s_ = 0.0;                                                       \
            HOST_THREAD_IDS(me_, all_);                                            \
            lo_ = (long int)((N) * (double)me_ / all_);                            \
            hi_ = (long int)((N) * (double)(me_ + 1) / all_);                      \
            if (me_ == all_ - 1) hi_ = (N);                                        \
            for (i = lo_; i < hi_; i++) s_ += (TERM);                              \
            part_[me_] = s_;                                                       \
            if (me_ == 0) used_ = all_;                                            \
        }                                                                          \
        (out) = 0.0;                                                               \
        for (t_ = 0; t_ < used_; t_++) (out) += part_[t_];                         \
    } while (0)
#if defined(_OPENMP)
#define HOST_THREAD_IDS(me, all) ((me) = omp_get_thread_num(), (all) = omp_get_num_threads())
#else
#define HOST_THREAD_IDS(me, all) ((me) = 0, (all) = 1)
#endif

/* N_VDotProd_Serial :637 (left to right) */
static double host_dot(N_Vector x, N_Vector y)
{
    long int N = HL(x); double *xd = HD(x), *yd = HD(y), sum; int nt = NT(x);
    HOST_BLOCK_SUM(N, nt, xd[i] * yd[i], sum);
    return sum;
}
static double host_maxnorm(N_Vector x)
{
    long int i, N = HL(x); double *xd = HD(x), mx = 0.0; int nt = NT(x); (void)nt;
#pragma omp parallel for reduction(max:mx) num_threads(nt) if (nt > 1)
    for (i = 0; i < N; i++) if (fabs(xd[i]) > mx) mx = fabs(xd[i]);
    return mx;
}
/* N_VWrmsNorm_Serial :673-691 */
static double host_wrms(N_Vector x, N_Vector w)
{
    long int N = HL(x); double *xd = HD(x), *wd = HD(w), sum; int nt = NT(x);
    HOST_BLOCK_SUM(N, nt, (xd[i] * wd[i]) * (xd[i] * wd[i]), sum);
    return sqrt(sum / N);
}
static double host_wrmsmask(N_Vector x, N_Vector w, N_Vector id)
{
    long int i, N = HL(x); double *xd = HD(x), *wd = HD(w), *idd = HD(id), sum = 0.0;
    for (i = 0; i < N; i++)
        if (idd[i] > 0.0) { double p = xd[i] * wd[i]; sum += p * p; }
    return sqrt(sum / N);
}
static double host_min(N_Vector x)
{
    long int i, N = HL(x); double *xd = HD(x), mn = xd[0]; int nt = NT(x); (void)nt;
#pragma omp parallel for reduction(min:mn) num_threads(nt) if (nt > 1)
    for (i = 1; i < N; i++) if (xd[i] < mn) mn = xd[i];
    return mn;
}
static double host_wl2(N_Vector x, N_Vector w)
{
    long int i, N = HL(x); double *xd = HD(x), *wd = HD(w), sum = 0.0;
    for (i = 0; i < N; i++) { double p = xd[i] * wd[i]; sum += p * p; }
    return sqrt(sum);
}
static double host_l1(N_Vector x)
{
    long int i, N = HL(x); double *xd = HD(x), sum = 0.0;
    for (i = 0; i < N; i++) sum += fabs(xd[i]);
    return sum;
}
static void host_compare(double c, N_Vector x, N_Vector z)
{
    long int i, N = HL(x); double *xd = HD(x), *zd = HD(z);
    for (i = 0; i < N; i++) zd[i] = (fabs(xd[i]) >= c) ? 1.0 : 0.0;
}
static booleantype host_invtest(N_Vector x, N_Vector z)
{
    long int i, N = HL(x); double *xd = HD(x), *zd = HD(z); int ok = 1;
    for (i = 0; i < N; i++) { if (xd[i] == 0.0) ok = 0; else zd[i] = 1.0 / xd[i]; }
    return ok;
}
static booleantype host_constrmask(N_Vector c, N_Vector x, N_Vector m)
{
    long int i, N = HL(x); double *cd = HD(c), *xd = HD(x), *md = HD(m); int test = 1;
    for (i = 0; i < N; i++) {
        md[i] = 0.0;
        if (cd[i] == 0.0) continue;
        if (cd[i] > 1.5 || cd[i] < -1.5) {
            if (xd[i] * cd[i] <= 0.0) { test = 0; md[i] = 1.0; }
            continue;
        }
        if (cd[i] > 0.5 || cd[i] < -0.5)
            if (xd[i] * cd[i] < 0.0) { test = 0; md[i] = 1.0; }
    }
    return test;
}
static double host_minquotient(N_Vector num, N_Vector den)
{
    long int i, N = HL(num); double *nd = HD(num), *dd = HD(den), mn = DBL_MAX; int first = 1;
    for (i = 0; i < N; i++) {
        if (dd[i] == 0.0) continue;
        if (!first) { double q = nd[i] / dd[i]; mn = (q < mn) ? q : mn; }
        else { mn = nd[i] / dd[i]; first = 0; }
    }
    return mn;
}

static struct _generic_N_Vector_Ops host_ops = {
    host_id, host_clone, host_cloneempty, oracle_nv_destroy, host_space,
    host_getarray, host_setarray, host_linearsum, host_const, host_prod,
    host_div, host_scale, host_abs, host_inv, host_addconst, host_dot,
    host_maxnorm, host_wrms, host_wrmsmask, host_min, host_wl2, host_l1,
    host_compare, host_invtest, host_constrmask, host_minquotient};

N_Vector oracle_nv_new(long int length, int nthreads)
{
    N_Vector v = (N_Vector)malloc(sizeof(*v));
    host_content *c = (host_content *)malloc(sizeof(*c));
    c->length = length; c->own_data = 1; c->nthreads = nthreads > 0 ? nthreads : 1;
    c->data = (double *)calloc(length > 0 ? (size_t)length : 1, sizeof(double));
    v->content = c; v->ops = &host_ops;
    return v;
}
static N_Vector host_cloneempty(N_Vector w)
{
    N_Vector v = (N_Vector)malloc(sizeof(*v));
    host_content *c = (host_content *)malloc(sizeof(*c));
    c->length = HL(w); c->own_data = 0; c->data = NULL; c->nthreads = NT(w);
    v->content = c; v->ops = &host_ops;
    return v;
}
static N_Vector host_clone(N_Vector w)
{
    N_Vector v = host_cloneempty(w);
    HC(v)->data = (double *)calloc(HL(w) > 0 ? (size_t)HL(w) : 1, sizeof(double));
    HC(v)->own_data = 1;
    return v;
}
void oracle_nv_destroy(N_Vector v)
{
    if (!v) return;
    if (HC(v)->own_data) free(HC(v)->data);
    free(v->content); free(v);
}
double *oracle_nv_data(N_Vector v) { return HD(v); }
long int oracle_nv_length(N_Vector v) { return HL(v); }
