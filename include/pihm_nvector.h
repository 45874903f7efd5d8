/*
 * pihm_nvector.h - device-resident N_Vector for MI355X (gfx950).
 *
 * The vector object has the memory layout of the SUNDIALS 2.7.0 generic
 * N_Vector (reference: cvode/include/sundials/sundials_nvector.h:90-130):
 * a `content` pointer followed by a pointer to a table of 26 operations in
 * the same order.  CVODE 2.9 (and the integrator in this library) call the
 * operations only through that table, so a vector made by N_VNew_Hip can be
 * handed to either.  If the SUNDIALS header was included first, its
 * definitions are used and the ones below are skipped.
 *
 * Data lives in device memory (HBM).  Reductions (dot product, norms, min)
 * are deterministic: a fixed-shape two-pass tree (wavefront shuffles inside a
 * block, then one block over the block partials in index order), so the same
 * input gives the same bits on every call and for every launch timing.  They
 * return a host double, which implies a stream synchronisation.
 */
#ifndef PIHM_NVECTOR_H
#define PIHM_NVECTOR_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef _SUNDIALSTYPES_H
typedef double realtype;
#ifndef booleantype
#define booleantype int
#endif
#endif

#ifndef _NVECTOR_H
/* Same enumerators and order as sundials_nvector.h:64-72. */
typedef enum {
  SUNDIALS_NVEC_SERIAL,
  SUNDIALS_NVEC_PARALLEL,
  SUNDIALS_NVEC_OPENMP,
  SUNDIALS_NVEC_PTHREADS,
  SUNDIALS_NVEC_PARHYP,
  SUNDIALS_NVEC_PETSC,
  SUNDIALS_NVEC_CUSTOM
} N_Vector_ID;

typedef struct _generic_N_Vector_Ops *N_Vector_Ops;
typedef struct _generic_N_Vector *N_Vector;
typedef N_Vector *N_Vector_S;

/* Operation table; order and signatures follow sundials_nvector.h:90-125. */
struct _generic_N_Vector_Ops {
  N_Vector_ID (*nvgetvectorid)(N_Vector);
  N_Vector    (*nvclone)(N_Vector);
  N_Vector    (*nvcloneempty)(N_Vector);
  void        (*nvdestroy)(N_Vector);
  void        (*nvspace)(N_Vector, long int *, long int *);
  realtype*   (*nvgetarraypointer)(N_Vector);
  void        (*nvsetarraypointer)(realtype *, N_Vector);
  void        (*nvlinearsum)(realtype, N_Vector, realtype, N_Vector, N_Vector);
  void        (*nvconst)(realtype, N_Vector);
  void        (*nvprod)(N_Vector, N_Vector, N_Vector);
  void        (*nvdiv)(N_Vector, N_Vector, N_Vector);
  void        (*nvscale)(realtype, N_Vector, N_Vector);
  void        (*nvabs)(N_Vector, N_Vector);
  void        (*nvinv)(N_Vector, N_Vector);
  void        (*nvaddconst)(N_Vector, realtype, N_Vector);
  realtype    (*nvdotprod)(N_Vector, N_Vector);
  realtype    (*nvmaxnorm)(N_Vector);
  realtype    (*nvwrmsnorm)(N_Vector, N_Vector);
  realtype    (*nvwrmsnormmask)(N_Vector, N_Vector, N_Vector);
  realtype    (*nvmin)(N_Vector);
  realtype    (*nvwl2norm)(N_Vector, N_Vector);
  realtype    (*nvl1norm)(N_Vector);
  void        (*nvcompare)(realtype, N_Vector, N_Vector);
  booleantype (*nvinvtest)(N_Vector, N_Vector);
  booleantype (*nvconstrmask)(N_Vector, N_Vector, N_Vector);
  realtype    (*nvminquotient)(N_Vector, N_Vector);
};

struct _generic_N_Vector {
  void *content;
  struct _generic_N_Vector_Ops *ops;
};
#endif /* _NVECTOR_H */

/* Vector id reported by nvgetvectorid for device vectors. */
#define PIHM_NVEC_HIP SUNDIALS_NVEC_CUSTOM

/* Content of a device vector.  `data` is a device pointer. */
struct _N_VectorContent_Hip {
  long int length;
  int      own_data;
  double  *data;
  void    *stream;      /* hipStream_t the operations are issued on */
  void    *aux;         /* library-owned reduction workspace */
};
typedef struct _N_VectorContent_Hip *N_VectorContent_Hip;

/* Constructors / destructor.
 * Replaces N_VNew_Serial (cvode/src/nvec_ser/nvector_serial.c) for device
 * data; the reference's CUDA build bound N_VNew_Cuda (src/main.c:74-75). */
N_Vector N_VNew_Hip(long int length);
N_Vector N_VNewEmpty_Hip(long int length);
N_Vector N_VMake_Hip(long int length, double *device_data);
void     N_VDestroy_Hip(N_Vector v);

/* Stream the vector's operations run on (NULL = the null stream). */
void     N_VSetStream_Hip(N_Vector v, void *hip_stream);

/* Host <-> device copies (the reference's CUDA build used
 * N_VCopyToDevice_Cuda / N_VCopyFromDevice_Cuda, src/util_func.c:205-222).
 * They copy `length` doubles and synchronise the vector's stream. */
int      N_VCopyToDevice_Hip(N_Vector v, const double *host);
int      N_VCopyFromDevice_Hip(N_Vector v, double *host);
double  *N_VGetDeviceArrayPointer_Hip(N_Vector v);
long int N_VGetLength_Hip(N_Vector v);

/* Operations of the table, callable directly. */
N_Vector_ID N_VGetVectorID_Hip(N_Vector v);
N_Vector N_VClone_Hip(N_Vector w);
N_Vector N_VCloneEmpty_Hip(N_Vector w);
void     N_VSpace_Hip(N_Vector v, long int *lrw, long int *liw);
double  *N_VGetArrayPointer_Hip(N_Vector v);
void     N_VSetArrayPointer_Hip(double *v_data, N_Vector v);
void     N_VLinearSum_Hip(double a, N_Vector x, double b, N_Vector y, N_Vector z);
void     N_VConst_Hip(double c, N_Vector z);
void     N_VProd_Hip(N_Vector x, N_Vector y, N_Vector z);
void     N_VDiv_Hip(N_Vector x, N_Vector y, N_Vector z);
void     N_VScale_Hip(double c, N_Vector x, N_Vector z);
void     N_VAbs_Hip(N_Vector x, N_Vector z);
void     N_VInv_Hip(N_Vector x, N_Vector z);
void     N_VAddConst_Hip(N_Vector x, double b, N_Vector z);
double   N_VDotProd_Hip(N_Vector x, N_Vector y);
double   N_VMaxNorm_Hip(N_Vector x);
double   N_VWrmsNorm_Hip(N_Vector x, N_Vector w);
double   N_VWrmsNormMask_Hip(N_Vector x, N_Vector w, N_Vector id);
double   N_VMin_Hip(N_Vector x);
double   N_VWL2Norm_Hip(N_Vector x, N_Vector w);
double   N_VL1Norm_Hip(N_Vector x);
void     N_VCompare_Hip(double c, N_Vector x, N_Vector z);
booleantype N_VInvTest_Hip(N_Vector x, N_Vector z);
booleantype N_VConstrMask_Hip(N_Vector c, N_Vector x, N_Vector m);
double   N_VMinQuotient_Hip(N_Vector num, N_Vector denom);

/* Fused operations (no counterpart in SUNDIALS 2.7; SUNDIALS 4 added the
 * same shapes as N_VLinearCombination / N_VDotProdMulti).  One launch each.
 *   z = sum_k c[k] * X[k]                      (nvec <= 16)
 *   d[k] = <x, Y[k]>                           (nvec <= 16, one sync)
 *   z = a*x + b*y  and return WRMS(z, w)       (one launch + one sync) */
int      N_VLinearCombination_Hip(int nvec, const double *c, N_Vector *X, N_Vector z);
int      N_VDotProdMulti_Hip(int nvec, N_Vector x, N_Vector *Y, double *d);
double   N_VLinearSumWrms_Hip(double a, N_Vector x, double b, N_Vector y,
                              N_Vector z, N_Vector w);

/* Measurement, no counterpart in the reference: average device time [ms] of
 * N_VLinearSum_Hip (z = 1.5 x + 0.5 y: 24 bytes per entry) and of
 * N_VDotProd_Hip (its two passes and the host's wait for the result) on new
 * vectors of length n, each over `iters` calls between HIP events on the
 * vectors' stream.  Returns 0, -1 for bad arguments, -2 on a runtime error. */
int      pihm_nv_time_ops(long int n, int iters, double *ms_linearsum, double *ms_dotprod);

/* No counterpart in the reference or in SUNDIALS.  For callers that cannot
 * allocate device or pinned memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): N_VDotProd_Hip on raw
 * device arrays, *result_dev = <x, y>, launched on hip_stream (NULL = the null
 * stream) without waiting.  Same summation tree, so the same bits as
 * N_VDotProd_Hip.  The workspace is a fixed 8 KiB array in the library's
 * device memory (each GPU holds its own until the process ends), so no call
 * allocates; calls that overlap on one GPU share it and race.  Returns 0,
 * -1 for n < 0, n > PIHM_NOALLOC_MAX_LENGTH or a NULL pointer, -2 on a
 * runtime error. */
#define PIHM_NOALLOC_MAX_LENGTH (1L << 33)
int      pihm_nv_dotprod_noalloc(long int n, const double *x_dev, const double *y_dev,
                                 double *result_dev, void *hip_stream);

#ifdef __cplusplus
}
#endif
#endif /* PIHM_NVECTOR_H */
