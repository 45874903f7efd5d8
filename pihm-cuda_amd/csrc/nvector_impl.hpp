// nvector_impl.hpp - internal interface of the device N_Vector.
#pragma once
#include "../../include/pihm_nvector.h"

#ifdef __cplusplus
extern "C" {
#endif

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
