// nvector_impl.hpp - internal interface of the device N_Vector.
#pragma once
#include "../../include/pihm_nvector.h"

#ifdef __cplusplus
extern "C" {
#endif

// Reduction workspace shared by a family of device vectors (clones share it).
// The second pass of every reduction writes its results to `result` and then
// the running count of second passes to `flag`; the host waits for that count
// (pihm_red_wait) instead of synchronising the stream.
typedef struct pihm_redwork {
  double *partials;          // device, [PIHM_RED_MAXBLOCKS * PIHM_RED_MAXSUMS]
  double *result;            // pinned host (coherent), [PIHM_RED_MAXSUMS]
  unsigned *ctr;             // device: second passes completed
  volatile unsigned *flag;   // pinned host (coherent): copy of *ctr
  unsigned *expect;          // host: second passes launched and awaited
} pihm_redwork;
#define PIHM_RED_MAXBLOCKS 1024
#define PIHM_RED_MAXSUMS 16
pihm_redwork pihm_nv_redwork(N_Vector v);
void *pihm_nv_stream(N_Vector v);
// Wait until the most recently launched second pass of w has written its
// results (spins on the pinned flag; falls back to synchronising the stream).
int pihm_red_wait(pihm_redwork w, void *stream);

// Fused Newton / SPGMR operations (krylov_hip.hip).  All vectors are device
// vectors of one family; s is the error-weight vector (the SPGMR scalings
// s1 = s2 = ewt of cvode_spgmr.c:396-398).  Sums are returned through the
// family's pinned result buffer after one stream synchronisation.
//   start:   b = gamma*f - (rl1*zn1 + acor); v0 = s*b;  returns sum (s*b)^2
//   prep:    V_l <- c * (V_l - sum_{k<l} coef[k] V_k)  (in place, l>0),
//            V_0 <- c * V_0 (l=0);  work = sig*(V_l/s) + y
//   arnoldi: w = s*((-gamma*(siginv*(jf - fy))) + V_l/s) into V_{l+1};
//            sums[k] = <V_k, w> (k<=l), sums[l+1] = <w, w>
//   orth:    V_{l+1} <- V_{l+1} - sum_{k<=l} coef[k] V_k; sums as arnoldi
//   finish:  x = (sum_k ycoef[k] V_k)/s (mode 0) | b (mode 1) | 0 (mode 2);
//            acor += x; y = zn0 + acor; sums: (x*s)^2, (acor*s)^2
//
// kry (optional, device, PIHM_KRY_SIZE doubles) receives a device copy of the
// sums: start -> kry[PIHM_KRY_START], arnoldi/orth step l -> row l
// (kry + l*PIHM_KRY_LD).  prep_dev forms V_l and the work vector with the
// coefficients of row l-1 (or the start sum for l = 0) read from kry, as the
// host computes them, and arnoldi with sums == NULL only launches, reading
// gamma from kry[PIHM_KRY_GAMMA] (stored there by start): a Krylov iteration
// can then be captured in a graph and replayed with no host arguments.
#define PIHM_KRY_LD 12
#define PIHM_KRY_START (11 * PIHM_KRY_LD)
#define PIHM_KRY_GAMMA (PIHM_KRY_START + 1)
#define PIHM_KRY_SIZE (PIHM_KRY_START + 4)
int pihm_gm_start(double rl1, N_Vector zn1, N_Vector acor, double gamma, N_Vector f,
                  N_Vector s, N_Vector b, N_Vector v0, double *sumsq, double *kry);
// With right preconditioning: prep writes V_l / s into vt (and not the work
// vector), the caller applies P^-1 and forms the RHS argument, and arnoldi
// takes the preconditioned vector z for the A*z step.  finish mode 3 takes the
// correction x from xsrc; combine forms sum_j y_j V_j / s for it.
int pihm_gm_prep(int l, N_Vector *V, const double *coef, double c, N_Vector y, N_Vector s,
                 double sig, N_Vector work, N_Vector vt);
int pihm_gm_prep_dev(int l, N_Vector *V, const double *kry, N_Vector y, N_Vector s, double sig,
                     N_Vector work, N_Vector vt);
int pihm_gm_arnoldi(int l, N_Vector *V, N_Vector s, N_Vector jf, N_Vector fy, double siginv,
                    double gamma, double *sums, double *kry, N_Vector z);
int pihm_gm_orth(int l, N_Vector *V, const double *coef, double *sums, double *kry);
int pihm_gm_finish(int mode, int k, const double *ycoef, N_Vector *V, N_Vector s, N_Vector b,
                   N_Vector zn0, N_Vector acor, N_Vector y, const double *xsrc, double *sums);
int pihm_gm_combine(int k, const double *ycoef, N_Vector *V, N_Vector s, N_Vector out);

// Fused Nordsieck-array updates used by the integrator (one launch each).
// zn holds q+1 (or more) device vectors of equal length.
//   predict:  for k=1..q, j=q..k: zn[j-1] += zn[j]        (cvode.c:2277-2306)
//   restore:  for k=1..q, j=q..k: zn[j-1] -= zn[j]        (cvode.c:2882-2913)
//   rescale:  zn[j] *= eta^j, j=1..q                      (cvode.c:2251-2275)
//   correct:  zn[j] += l[j]*acor, j=0..q                  (cvode.c:3010-3011)
int pihm_nv_predict(int q, N_Vector *zn);
int pihm_nv_restore(int q, N_Vector *zn);
int pihm_nv_rescale(int q, const double *factors, N_Vector *zn);
int pihm_nv_correct(int q, const double *l, N_Vector acor, N_Vector *zn);
// ewt = 1/(rtol*|y| + atol); returns min(rtol*|y|+atol) in *mn and
// WRMS(y, ewt) in *ynrm (cvode.c:4081-4098 and :1377-1378) in one pass.
int pihm_nv_ewt_ss(N_Vector y, double rtol, double atol, N_Vector ewt, double *mn, double *ynrm);
// Newton residual: tmp = gamma*f - (rl1*zn1 + acor)  (cvode.c:2744-2745)
int pihm_nv_newton_rhs(double rl1, N_Vector zn1, N_Vector acor, double gamma, N_Vector f,
                       N_Vector out);

#ifdef __cplusplus
}
#endif
