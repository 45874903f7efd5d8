/*
 * pihm_cvode.h - variable-order BDF integrator with Newton/SPGMR iteration.
 *
 * A CVODE-shaped C API covering what MM-PIHM uses (src/ode.c:413-661,
 * src/main.c:83-116): create, init/reinit, scalar tolerances, user data,
 * initial/max step, max steps, stop time, SPGMR without preconditioning
 * (maxl default 5) or with a right preconditioner, CV_NORMAL / CV_ONE_STEP solves and the counters read
 * by AdjCVodeMaxStep and PrintPerf.  Method: fixed-leading-coefficient BDF
 * of order 1..5 in Nordsieck form with the error test, step/order selection,
 * Newton iteration and difference-quotient Jacobian-vector products of
 * CVODE 2.9 (Hindmarsh et al., ACM TOMS 31:363-396, 2005; Brown, Byrne and
 * Hindmarsh, SIAM J. Sci. Stat. Comput. 10:1038-1051, 1989).  Stability
 * limit detection is accepted as an option and not performed.
 *
 * It works on any N_Vector whose operation table is filled (the host vector
 * of the test oracle, or the device vector of pihm_nvector.h); with device
 * vectors, groups of operations are fused into single launches.
 */
#ifndef PIHM_CVODE_H
#define PIHM_CVODE_H

#include "pihm_nvector.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Same values as cvode/include/cvode/cvode.h:82-126. */
#define PIHM_CV_BDF    2
#define PIHM_CV_NEWTON 2
#define PIHM_CV_NORMAL   1
#define PIHM_CV_ONE_STEP 2
#define PIHM_CV_SUCCESS        0
#define PIHM_CV_TSTOP_RETURN   1
#define PIHM_CV_TOO_MUCH_WORK -1
#define PIHM_CV_TOO_MUCH_ACC  -2
#define PIHM_CV_ERR_FAILURE   -3
#define PIHM_CV_CONV_FAILURE  -4
#define PIHM_CV_LSETUP_FAIL   -6
#define PIHM_CV_LSOLVE_FAIL   -7
#define PIHM_CV_RHSFUNC_FAIL  -8
#define PIHM_CV_FIRST_RHSFUNC_ERR -9
#define PIHM_CV_REPTD_RHSFUNC_ERR -10
#define PIHM_CV_UNREC_RHSFUNC_ERR -11
#define PIHM_CV_MEM_FAIL      -20
#define PIHM_CV_MEM_NULL      -21
#define PIHM_CV_ILL_INPUT     -22
#define PIHM_CV_NO_MALLOC     -23
#define PIHM_CV_BAD_K         -24
#define PIHM_CV_BAD_T         -25

typedef int (*PihmRhsFn)(realtype t, N_Vector y, N_Vector ydot, void *user_data);

/* Preconditioning of SPGMR: the values of sundials_iterative.h:57 and the
 * callback shapes of CVSpilsPrecSetupFn / CVSpilsPrecSolveFn
 * (cvode/include/cvode/cvode_spils.h:162-222).  Only right preconditioning
 * is provided. */
#define PIHM_PREC_NONE  0
#define PIHM_PREC_RIGHT 2
typedef int (*PihmPrecSetupFn)(realtype t, N_Vector y, N_Vector fy, int jok, int *jcurPtr, realtype gamma,
                               void *user_data, N_Vector tmp1, N_Vector tmp2, N_Vector tmp3);
typedef int (*PihmPrecSolveFn)(realtype t, N_Vector y, N_Vector fy, N_Vector r, N_Vector z, realtype gamma,
                               realtype delta, int lr, void *user_data, N_Vector tmp);

void *PihmCVodeCreate(int lmm, int iter);
int   PihmCVodeInit(void *mem, PihmRhsFn f, realtype t0, N_Vector y0);
int   PihmCVodeReInit(void *mem, realtype t0, N_Vector y0);
int   PihmCVodeSStolerances(void *mem, realtype reltol, realtype abstol);
int   PihmCVodeSetUserData(void *mem, void *user_data);
int   PihmCVodeSetInitStep(void *mem, realtype hin);
int   PihmCVodeSetMaxStep(void *mem, realtype hmax);
int   PihmCVodeSetMinStep(void *mem, realtype hmin);
int   PihmCVodeSetMaxNumSteps(void *mem, long int mxsteps);
int   PihmCVodeSetMaxOrd(void *mem, int maxord);
int   PihmCVodeSetStopTime(void *mem, realtype tstop);
int   PihmCVodeSetStabLimDet(void *mem, int onoff);
/* Declares that the RHS only enqueues device work on the vectors' stream (no
 * host-side effects, no synchronisation, result independent of t), as
 * PIHM_ODE_Hip with the NaN check off does.  The integrator then captures
 * each Krylov iteration (basis vector, RHS, Arnoldi step) once in a HIP graph
 * and replays it; device vectors on a non-null stream are needed, otherwise
 * (or if capture fails) it runs the same steps launch by launch.  Results are
 * the same bits either way.  No counterpart in CVODE. */
int   PihmCVodeSetGraphRhs(void *mem, int onoff);
/* pretype PIHM_PREC_NONE (MM-PIHM's choice, src/ode.c:523) or PIHM_PREC_RIGHT. */
int   PihmCVSpgmr(void *mem, int pretype, int maxl);
/* CVSpilsSetPreconditioner (cvode_spils.h): the callbacks get the user data.
 * With right preconditioning the setup runs when CVODE would call lsetup
 * (cvode.c:2651-2727, cvode_spgmr.c:289-330).  With device vectors the
 * difference-quotient increment of J*P^-1*v is the unpreconditioned one
 * (sqrt(N) for the unit basis vector), and the preconditioner solve must only
 * enqueue device work on the vectors' stream when graphs are on. */
int   PihmCVSpilsSetPreconditioner(void *mem, PihmPrecSetupFn pset, PihmPrecSolveFn psolve);
/* Step record and replay; no counterpart in CVODE.
 * Record: the size and order of every accepted step k go to h[k], q[k]
 * (k < cap), *n counting them.  With iters non-null, iters[k*PIHM_REC_ITERS]
 * holds the step's Newton iteration count and iters[k*PIHM_REC_ITERS + 1 + j]
 * the Krylov iterations of Newton iteration j (0: the residual was already
 * below the linear tolerance; -1: SPGMR returned the zero solution).
 * Replay: step k is taken with h[k] and q[k] (k < n) in place of the step and
 * order selection, and its error test is not applied; with iters non-null it
 * also takes exactly the recorded Newton and Krylov counts in place of their
 * convergence tests.  A Newton failure on a replayed step returns
 * PIHM_CV_CONV_FAILURE.  Past step n the normal selection resumes. */
#define PIHM_REC_ITERS 8
int   PihmCVodeSetStepRecord(void *mem, double *h, int *q, int *iters, long cap, long *n);
int   PihmCVodeSetStepReplay(void *mem, const double *h, const int *q, const int *iters, long n);
int   PihmCVode(void *mem, realtype tout, N_Vector yout, realtype *tret, int itask);
int   PihmCVodeGetDky(void *mem, realtype t, int k, N_Vector dky);
void  PihmCVodeFree(void **mem);

int PihmCVodeGetNumSteps(void *mem, long int *nsteps);
int PihmCVodeGetNumRhsEvals(void *mem, long int *nfevals);
int PihmCVodeGetNumLinIters(void *mem, long int *nliters);
int PihmCVodeGetNumJtimesRhsEvals(void *mem, long int *nfevals);
int PihmCVodeGetNumNonlinSolvIters(void *mem, long int *nniters);
int PihmCVodeGetNumNonlinSolvConvFails(void *mem, long int *nncfails);
int PihmCVodeGetNumErrTestFails(void *mem, long int *netfails);
int PihmCVodeGetLastStep(void *mem, realtype *hlast);
int PihmCVodeGetLastOrder(void *mem, int *qlast);
int PihmCVodeGetCurrentTime(void *mem, realtype *tcur);
int PihmCVodeGetCurrentStep(void *mem, realtype *hcur);
int PihmCVodeGetNumLinSolvSetups(void *mem, long int *nlinsetups);
int PihmCVSpilsGetNumPrecEvals(void *mem, long int *npevals);
int PihmCVSpilsGetNumPrecSolves(void *mem, long int *npsolves);

#ifdef __cplusplus
}
#endif
#endif
