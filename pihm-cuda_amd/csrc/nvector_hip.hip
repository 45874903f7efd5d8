// nvector_hip.hip - device-resident N_Vector (SUNDIALS 2.7 operation table).
//
// Element-wise operations follow the rounding of nvector_serial.c
// (cvode/src/nvec_ser/nvector_serial.c): N_VLinearSum evaluates a*(x+y) when
// a == b and a*(x-y) when a == -b, (a*x)+(b*y) otherwise, which equals every
// special case of the serial version bit for bit.
//
// Reductions run in two launches: each workgroup reduces a fixed, strided set
// of entries (wavefront shuffles, then LDS) and writes its partial; one
// workgroup then adds the partials in workgroup order and writes the result
// to pinned host memory.  The summation tree depends
// only on the length, so results are reproducible bit for bit.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "common.hpp"
#include "nvector_impl.hpp"
#include "redux.hpp"

using namespace pihm;

namespace {

constexpr int kRedBlock = 256;
constexpr int kMaxRedBlocks = PIHM_RED_MAXBLOCKS;
constexpr int kMaxMulti = PIHM_RED_MAXSUMS;

using RedWork = pihm_redwork;

// One workspace per device vector family (clones share the parent's).
struct Aux {
  RedWork w;
  int refs;
};

inline N_VectorContent_Hip C(N_Vector v) { return (N_VectorContent_Hip)v->content; }
inline double *D(N_Vector v) { return C(v)->data; }
inline long L(N_Vector v) { return C(v)->length; }
inline hipStream_t S(N_Vector v) { return (hipStream_t)C(v)->stream; }
inline Aux *X(N_Vector v) { return (Aux *)C(v)->aux; }

inline int ew_blocks(long n) { return (int)std::max<long>(1, (n + kBlock - 1) / kBlock); }
inline int red_blocks(long n) {
  long b = (n + 2L * kRedBlock - 1) / (2L * kRedBlock);
  return (int)std::min<long>(std::max<long>(b, 1), kMaxRedBlocks);
}

// ---------------------------------------------------------------- kernels
__global__ void k_linsum(long n, double a, const double *x, double b, const double *y, double *z,
                         int mode) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xv = x[i], yv = y[i];
  double r;
  if (mode == 1) r = a * (xv + yv);
  else if (mode == 2) r = a * (xv - yv);
  else r = (a * xv) + (b * yv);
  z[i] = r;
}
__global__ void k_const(long n, double c, double *z) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = c;
}
// op: 0 prod, 1 div, 2 scale, 3 abs, 4 inv, 5 addconst, 6 compare, 7 neg/copy via scale
__global__ void k_unary(long n, int op, double c, const double *x, const double *y, double *z) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xv = x[i];
  double r;
  switch (op) {
    case 0: r = xv * y[i]; break;
    case 1: r = xv / y[i]; break;
    case 2: r = c * xv; break;
    case 3: r = fabs(xv); break;
    case 4: r = 1.0 / xv; break;
    case 5: r = xv + c; break;
    case 6: r = (fabs(xv) >= c) ? 1.0 : 0.0; break;
    case 7: r = -xv; break;
    default: r = xv; break;
  }
  z[i] = r;
}

// reduction kinds
enum RedKind { R_DOT = 0, R_WSQ, R_WSQMASK, R_MAXABS, R_MIN, R_ABSSUM, R_MINQ, R_INVTEST, R_CONSTR };

__device__ __forceinline__ double red_identity(int kind) {
  switch (kind) {
    case R_MAXABS: return 0.0;
    case R_MIN: case R_MINQ: return DBL_MAX;
    default: return 0.0;
  }
}
__device__ __forceinline__ double red_combine(int kind, double a, double b) {
  switch (kind) {
    case R_MAXABS: return fmax(a, b);
    case R_MIN: case R_MINQ: return fmin(a, b);
    case R_INVTEST: case R_CONSTR: return fmax(a, b);
    default: return a + b;
  }
}

__device__ __forceinline__ double wave_reduce(int kind, double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = red_combine(kind, v, __shfl_xor(v, off, 64));
  return v;
}

// Block reduction in a fixed order: wavefront butterfly, then wave 0 over
// the 4 wave sums.
__device__ __forceinline__ double block_reduce(int kind, double v, double *lds) {
  v = wave_reduce(kind, v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double r = red_identity(kind);
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) r = red_combine(kind, r, lds[w]);
  }
  __syncthreads();
  return r;  // valid in thread 0
}

struct RedArgs {
  long n;
  int kind;
  int nvec;                    // number of simultaneous reductions (dot-multi)
  const double *x;
  const double *w;             // weights / second operand
  const double *id;            // mask / constraint
  double *m;                   // constrmask output
  double *z;                   // invtest output
  const double *ys[kMaxMulti]; // dot-multi operands
  RedWork work;
};

__device__ __forceinline__ double red_elem(const RedArgs &a, long i, int k) {
  switch (a.kind) {
    case R_DOT: return a.x[i] * a.ys[k][i];
    case R_WSQ: { double p = a.x[i] * a.w[i]; return p * p; }
    case R_WSQMASK: { if (a.id[i] > 0.0) { double p = a.x[i] * a.w[i]; return p * p; } return 0.0; }
    case R_MAXABS: return fabs(a.x[i]);
    case R_MIN: return a.x[i];
    case R_ABSSUM: return fabs(a.x[i]);
    case R_MINQ: { double d = a.w[i]; return (d == 0.0) ? DBL_MAX : a.x[i] / d; }
    case R_INVTEST: {
      double xv = a.x[i];
      if (xv == 0.0) return 1.0;
      a.z[i] = 1.0 / xv;
      return 0.0;
    }
    case R_CONSTR: {
      const double c = a.id[i], xv = a.x[i];
      double bad = 0.0;
      if (c != 0.0) {
        if (c > 1.5 || c < -1.5) { if (xv * c <= 0.0) bad = 1.0; }
        else if (c > 0.5 || c < -0.5) { if (xv * c < 0.0) bad = 1.0; }
      }
      a.m[i] = bad;
      return bad;
    }
  }
  return 0.0;
}

__global__ __launch_bounds__(kRedBlock) void k_reduce(RedArgs a) {
  __shared__ double lds[4];
  const long stride = (long)gridDim.x * blockDim.x;
  for (int k = 0; k < a.nvec; k++) {
    double v = red_identity(a.kind);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride)
      v = red_combine(a.kind, v, red_elem(a, i, k));
    const double r = block_reduce(a.kind, v, lds);
    if (threadIdx.x == 0) a.work.partials[(long)k * kMaxRedBlocks + blockIdx.x] = r;
  }
}

// how the second pass combines the partials of a reduction kind
unsigned red_comb(int kind) {
  switch (kind) {
    case R_MAXABS: case R_INVTEST: case R_CONSTR: return COMB_MAX;
    case R_MIN: case R_MINQ: return COMB_MIN;
    default: return COMB_SUM;
  }
}

// fused ops ------------------------------------------------------------------
struct ZnPtrs { double *p[7]; };

template <int Q>
__global__ void k_predict(long n, ZnPtrs z, int sign) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v[Q + 1];
#pragma unroll
  for (int j = 0; j <= Q; j++) v[j] = z.p[j][i];
#pragma unroll
  for (int k = 1; k <= Q; k++)
#pragma unroll
    for (int j = Q; j >= k; j--) {
      if (sign > 0) v[j - 1] = v[j - 1] + v[j];   // LinearSum(1, zn[j-1], 1, zn[j]) == x+y
      else v[j - 1] = v[j - 1] - v[j];            // LinearSum(1, zn[j-1], -1, zn[j]) == x-y
    }
#pragma unroll
  for (int j = 0; j < Q; j++) z.p[j][i] = v[j];
}

void launch_predict(long n, int q, ZnPtrs z, int sign, hipStream_t st) {
  const dim3 g(ew_blocks(n)), b(kBlock);
  switch (q) {
    case 1: hipLaunchKernelGGL(k_predict<1>, g, b, 0, st, n, z, sign); break;
    case 2: hipLaunchKernelGGL(k_predict<2>, g, b, 0, st, n, z, sign); break;
    case 3: hipLaunchKernelGGL(k_predict<3>, g, b, 0, st, n, z, sign); break;
    case 4: hipLaunchKernelGGL(k_predict<4>, g, b, 0, st, n, z, sign); break;
    default: hipLaunchKernelGGL(k_predict<5>, g, b, 0, st, n, z, sign); break;
  }
}

__global__ void k_rescale(long n, int q, ZnPtrs z, double f1, double f2, double f3, double f4, double f5) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double f[6] = {1.0, f1, f2, f3, f4, f5};
  // every entry loaded before any is stored (a store would keep the next
  // array's load behind it)
  double v[6];
#pragma unroll
  for (int j = 1; j < 6; j++) v[j] = (j <= q) ? z.p[j][i] : 0.0;
#pragma unroll
  for (int j = 1; j < 6; j++)
    if (j <= q) z.p[j][i] = f[j] * v[j];
}

__global__ void k_correct(long n, int q, ZnPtrs z, const double *acor, double l0, double l1, double l2,
                          double l3, double l4, double l5) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double l[6] = {l0, l1, l2, l3, l4, l5};
  const double a = acor[i];
  double v[6];
#pragma unroll
  for (int j = 0; j < 6; j++) v[j] = (j <= q) ? z.p[j][i] : 0.0;   // all loads before any store
#pragma unroll
  for (int j = 0; j < 6; j++) {
    if (j > q) break;
    // LinearSum(l[j], acor, 1, zn[j], zn[j]): axpy y += a*x (or += x, -= x)
    const double lj = l[j];
    double zj = v[j];
    if (lj == 1.0) zj += a;
    else if (lj == -1.0) zj -= a;
    else zj += lj * a;
    z.p[j][i] = zj;
  }
}

__global__ __launch_bounds__(kRedBlock) void k_ewt(long n, const double *y, double rtol, double atol,
                                                  double *ewt, RedWork work) {
  __shared__ double lds[4];
  const long stride = (long)gridDim.x * blockDim.x;
  double mn = DBL_MAX, sq = 0.0;
  // cvEwtSetSS: tempv = |y|; tempv = rtol*tempv; tempv += atol; ewt = 1/tempv
  const auto one = [&](long i, double yi) {
    const double t = (rtol * fabs(yi)) + atol;
    const double w = 1.0 / t;
    ewt[i] = w;
    mn = fmin(mn, t);
    const double p = yi * w;
    sq += p * p;
  };
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + stride < n; i += 2 * stride) {   // two entries' loads before either store
    const double y0 = y[i], y1 = y[i + stride];
    one(i, y0);
    one(i + stride, y1);
  }
  if (i < n) one(i, y[i]);
  const double r0 = block_reduce(R_MIN, mn, lds);
  const double r1 = block_reduce(R_DOT, sq, lds);
  if (threadIdx.x == 0) {
    work.partials[blockIdx.x] = r0;
    work.partials[kMaxRedBlocks + blockIdx.x] = r1;
  }
}

__global__ void k_newton_rhs(long n, double rl1, const double *zn1, const double *acor, double gamma,
                             const double *f, double *out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double t = (rl1 * zn1[i]) + acor[i];
  out[i] = (gamma * f[i]) - t;
}

struct LcArgs { const double *x[kMaxMulti]; double c[kMaxMulti]; };
__global__ void k_lincomb(long n, int nvec, LcArgs a, double *z) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = a.c[0] * a.x[0][i];
  for (int k = 1; k < nvec; k++) s = s + a.c[k] * a.x[k][i];
  z[i] = s;
}

// ---------------------------------------------------------------- helpers
// Pinned host block of a workspace: results, then the completion count.
struct HostBlock {
  double result[kMaxMulti];
  unsigned flag;
  unsigned expect;
};

int aux_new(Aux **out) {
  Aux *x = (Aux *)std::calloc(1, sizeof(Aux));
  if (!x) return -1;
  HostBlock *hb = nullptr;
  if (hipMalloc(&x->w.partials, sizeof(double) * kMaxRedBlocks * kMaxMulti) != hipSuccess) return -1;
  if (hipMalloc(&x->w.ctr, sizeof(unsigned)) != hipSuccess) return -1;
  if (hipMemset(x->w.ctr, 0, sizeof(unsigned)) != hipSuccess) return -1;
  if (hipHostMalloc((void **)&hb, sizeof(HostBlock), hipHostMallocCoherent) != hipSuccess) return -1;
  std::memset(hb, 0, sizeof(HostBlock));
  x->w.result = hb->result;
  x->w.flag = &hb->flag;
  x->w.expect = &hb->expect;
  x->refs = 1;
  *out = x;
  return 0;
}
void aux_release(Aux *x) {
  if (!x) return;
  if (--x->refs > 0) return;
  (void)hipFree(x->w.partials);
  (void)hipFree(x->w.ctr);
  (void)hipHostFree(x->w.result);   // first member of the host block
  std::free(x);
}

bool reduce(N_Vector v, RedArgs &a, double *out, int nout) {
  a.work = X(v)->w;
  const int nb = red_blocks(a.n);
  hipLaunchKernelGGL(k_reduce, dim3(nb), dim3(kRedBlock), 0, S(v), a);
  unsigned combs = 0;
  for (int k = 0; k < a.nvec; k++) combs |= red_comb(a.kind) << (2 * k);
  launch_final(nb, a.nvec, combs, a.work, S(v), nullptr);
  if (pihm_red_wait(a.work, S(v))) return false;
  for (int k = 0; k < nout; k++) out[k] = a.work.result[k];
  return true;
}

void launch_unary(N_Vector x, int op, double c, N_Vector y, N_Vector z) {
  const long n = L(x);
  hipLaunchKernelGGL(k_unary, dim3(ew_blocks(n)), dim3(kBlock), 0, S(x), n, op, c, D(x),
                     y ? D(y) : nullptr, D(z));
}

struct _generic_N_Vector_Ops *hip_ops();

N_Vector make_vec(long length, double *data, int own, void *stream, Aux *aux) {
  N_Vector v = (N_Vector)std::malloc(sizeof(*v));
  N_VectorContent_Hip c = (N_VectorContent_Hip)std::malloc(sizeof(*c));
  c->length = length;
  c->own_data = own;
  c->data = data;
  c->stream = stream;
  c->aux = aux;
  v->content = c;
  v->ops = hip_ops();
  return v;
}

}  // namespace

// ---------------------------------------------------------------- public API
extern "C" {

N_Vector_ID N_VGetVectorID_Hip(N_Vector v) { (void)v; return PIHM_NVEC_HIP; }

pihm_redwork pihm_nv_redwork(N_Vector v) { return X(v)->w; }

int pihm_red_wait(pihm_redwork w, void *stream) {
  const unsigned want = ++(*w.expect);
  // Spin on the count the second pass writes to host memory; after 2 ms hand
  // over to the runtime's wait, which also reports a failed launch.
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned it = 1; *w.flag != want; it++) {
    if ((it & 255u) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
      if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return -2;
      if (*w.flag != want) return -2;
      break;
    }
    __builtin_ia32_pause();
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return 0;
}
void *pihm_nv_stream(N_Vector v) { return C(v)->stream; }

N_Vector N_VNewEmpty_Hip(long int length) {
  Aux *aux = nullptr;
  if (aux_new(&aux)) return nullptr;
  return make_vec(length, nullptr, 0, nullptr, aux);
}

N_Vector N_VNew_Hip(long int length) {
  N_Vector v = N_VNewEmpty_Hip(length);
  if (!v) return nullptr;
  double *d = nullptr;
  if (hipMalloc(&d, sizeof(double) * std::max<long>(length, 1)) != hipSuccess) {
    N_VDestroy_Hip(v);
    return nullptr;
  }
  (void)hipMemset(d, 0, sizeof(double) * std::max<long>(length, 1));
  C(v)->data = d;
  C(v)->own_data = 1;
  return v;
}

N_Vector N_VMake_Hip(long int length, double *device_data) {
  N_Vector v = N_VNewEmpty_Hip(length);
  if (v) C(v)->data = device_data;
  return v;
}

void N_VDestroy_Hip(N_Vector v) {
  if (!v) return;
  if (C(v)->own_data && C(v)->data) (void)hipFree(C(v)->data);
  aux_release(X(v));
  std::free(v->content);
  std::free(v);
}

N_Vector N_VCloneEmpty_Hip(N_Vector w) {
  X(w)->refs++;
  return make_vec(L(w), nullptr, 0, C(w)->stream, X(w));
}

N_Vector N_VClone_Hip(N_Vector w) {
  N_Vector v = N_VCloneEmpty_Hip(w);
  double *d = nullptr;
  if (hipMalloc(&d, sizeof(double) * std::max<long>(L(w), 1)) != hipSuccess) {
    N_VDestroy_Hip(v);
    return nullptr;
  }
  C(v)->data = d;
  C(v)->own_data = 1;
  return v;
}

void N_VSetStream_Hip(N_Vector v, void *s) { C(v)->stream = s; }

void N_VSpace_Hip(N_Vector v, long int *lrw, long int *liw) { *lrw = L(v); *liw = 2; }
double *N_VGetArrayPointer_Hip(N_Vector v) { return D(v); }
double *N_VGetDeviceArrayPointer_Hip(N_Vector v) { return D(v); }
void N_VSetArrayPointer_Hip(double *d, N_Vector v) { if (L(v) > 0) C(v)->data = d; }
long int N_VGetLength_Hip(N_Vector v) { return L(v); }

int N_VCopyToDevice_Hip(N_Vector v, const double *host) {
  if (hipMemcpyAsync(D(v), host, sizeof(double) * L(v), hipMemcpyHostToDevice, S(v)) != hipSuccess) return -2;
  return hipStreamSynchronize(S(v)) == hipSuccess ? 0 : -2;
}
int N_VCopyFromDevice_Hip(N_Vector v, double *host) {
  if (hipMemcpyAsync(host, D(v), sizeof(double) * L(v), hipMemcpyDeviceToHost, S(v)) != hipSuccess) return -2;
  return hipStreamSynchronize(S(v)) == hipSuccess ? 0 : -2;
}

void N_VLinearSum_Hip(double a, N_Vector x, double b, N_Vector y, N_Vector z) {
  int mode = (a == b) ? 1 : ((a == -b) ? 2 : 0);
  const long n = L(x);
  hipLaunchKernelGGL(k_linsum, dim3(ew_blocks(n)), dim3(kBlock), 0, S(x), n, a, D(x), b, D(y), D(z), mode);
}
void N_VConst_Hip(double c, N_Vector z) {
  const long n = L(z);
  hipLaunchKernelGGL(k_const, dim3(ew_blocks(n)), dim3(kBlock), 0, S(z), n, c, D(z));
}
void N_VProd_Hip(N_Vector x, N_Vector y, N_Vector z) { launch_unary(x, 0, 0.0, y, z); }
void N_VDiv_Hip(N_Vector x, N_Vector y, N_Vector z) { launch_unary(x, 1, 0.0, y, z); }
void N_VScale_Hip(double c, N_Vector x, N_Vector z) {
  if (c == 1.0 && z != x) { launch_unary(x, 8, 0.0, nullptr, z); return; }
  if (c == -1.0 && z != x) { launch_unary(x, 7, 0.0, nullptr, z); return; }
  launch_unary(x, 2, c, nullptr, z);
}
void N_VAbs_Hip(N_Vector x, N_Vector z) { launch_unary(x, 3, 0.0, nullptr, z); }
void N_VInv_Hip(N_Vector x, N_Vector z) { launch_unary(x, 4, 0.0, nullptr, z); }
void N_VAddConst_Hip(N_Vector x, double b, N_Vector z) { launch_unary(x, 5, b, nullptr, z); }
void N_VCompare_Hip(double c, N_Vector x, N_Vector z) { launch_unary(x, 6, c, nullptr, z); }

static RedArgs red_init(N_Vector x, int kind) {
  RedArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = L(x);
  a.kind = kind;
  a.nvec = 1;
  a.x = D(x);
  return a;
}

double N_VDotProd_Hip(N_Vector x, N_Vector y) {
  RedArgs a = red_init(x, R_DOT);
  a.ys[0] = D(y);
  double r = NAN;
  reduce(x, a, &r, 1);
  return r;
}
double N_VMaxNorm_Hip(N_Vector x) {
  RedArgs a = red_init(x, R_MAXABS);
  double r = NAN;
  reduce(x, a, &r, 1);
  return r;
}
double N_VWrmsNorm_Hip(N_Vector x, N_Vector w) {
  RedArgs a = red_init(x, R_WSQ);
  a.w = D(w);
  double r = NAN;
  reduce(x, a, &r, 1);
  return std::sqrt(r / (double)L(x));
}
double N_VWrmsNormMask_Hip(N_Vector x, N_Vector w, N_Vector id) {
  RedArgs a = red_init(x, R_WSQMASK);
  a.w = D(w);
  a.id = D(id);
  double r = NAN;
  reduce(x, a, &r, 1);
  return std::sqrt(r / (double)L(x));
}
double N_VMin_Hip(N_Vector x) {
  RedArgs a = red_init(x, R_MIN);
  double r = NAN;
  reduce(x, a, &r, 1);
  return r;
}
double N_VWL2Norm_Hip(N_Vector x, N_Vector w) {
  RedArgs a = red_init(x, R_WSQ);
  a.w = D(w);
  double r = NAN;
  reduce(x, a, &r, 1);
  return std::sqrt(r);
}
double N_VL1Norm_Hip(N_Vector x) {
  RedArgs a = red_init(x, R_ABSSUM);
  double r = NAN;
  reduce(x, a, &r, 1);
  return r;
}
booleantype N_VInvTest_Hip(N_Vector x, N_Vector z) {
  RedArgs a = red_init(x, R_INVTEST);
  a.z = D(z);
  double r = 1.0;
  reduce(x, a, &r, 1);
  return r == 0.0;
}
booleantype N_VConstrMask_Hip(N_Vector c, N_Vector x, N_Vector m) {
  RedArgs a = red_init(x, R_CONSTR);
  a.id = D(c);
  a.m = D(m);
  double r = 1.0;
  reduce(x, a, &r, 1);
  return r == 0.0;
}
double N_VMinQuotient_Hip(N_Vector num, N_Vector denom) {
  RedArgs a = red_init(num, R_MINQ);
  a.w = D(denom);
  double r = DBL_MAX;
  reduce(num, a, &r, 1);
  return r;
}

int N_VLinearCombination_Hip(int nvec, const double *c, N_Vector *Xv, N_Vector z) {
  if (nvec < 1 || nvec > kMaxMulti) return -1;
  LcArgs a;
  for (int k = 0; k < nvec; k++) { a.x[k] = D(Xv[k]); a.c[k] = c[k]; }
  const long n = L(z);
  hipLaunchKernelGGL(k_lincomb, dim3(ew_blocks(n)), dim3(kBlock), 0, S(z), n, nvec, a, D(z));
  return 0;
}

int N_VDotProdMulti_Hip(int nvec, N_Vector x, N_Vector *Y, double *d) {
  if (nvec < 1 || nvec > kMaxMulti) return -1;
  RedArgs a = red_init(x, R_DOT);
  a.nvec = nvec;
  for (int k = 0; k < nvec; k++) a.ys[k] = D(Y[k]);
  return reduce(x, a, d, nvec) ? 0 : -2;
}

double N_VLinearSumWrms_Hip(double a, N_Vector x, double b, N_Vector y, N_Vector z, N_Vector w) {
  N_VLinearSum_Hip(a, x, b, y, z);
  return N_VWrmsNorm_Hip(z, w);
}

int pihm_nv_predict(int q, N_Vector *zn) {
  ZnPtrs z;
  for (int j = 0; j <= q; j++) z.p[j] = D(zn[j]);
  const long n = L(zn[0]);
  launch_predict(n, q, z, 1, S(zn[0]));
  return 0;
}
int pihm_nv_restore(int q, N_Vector *zn) {
  ZnPtrs z;
  for (int j = 0; j <= q; j++) z.p[j] = D(zn[j]);
  const long n = L(zn[0]);
  launch_predict(n, q, z, -1, S(zn[0]));
  return 0;
}
int pihm_nv_rescale(int q, const double *f, N_Vector *zn) {
  ZnPtrs z;
  for (int j = 0; j <= q; j++) z.p[j] = D(zn[j]);
  const long n = L(zn[0]);
  double ff[6] = {1, 1, 1, 1, 1, 1};
  for (int j = 1; j <= q && j <= 5; j++) ff[j] = f[j];
  hipLaunchKernelGGL(k_rescale, dim3(ew_blocks(n)), dim3(kBlock), 0, S(zn[0]), n, q, z, ff[1], ff[2],
                     ff[3], ff[4], ff[5]);
  return 0;
}
int pihm_nv_correct(int q, const double *l, N_Vector acor, N_Vector *zn) {
  ZnPtrs z;
  for (int j = 0; j <= q; j++) z.p[j] = D(zn[j]);
  const long n = L(zn[0]);
  double ll[6] = {0, 0, 0, 0, 0, 0};
  for (int j = 0; j <= q && j <= 5; j++) ll[j] = l[j];
  hipLaunchKernelGGL(k_correct, dim3(ew_blocks(n)), dim3(kBlock), 0, S(zn[0]), n, q, z, D(acor), ll[0],
                     ll[1], ll[2], ll[3], ll[4], ll[5]);
  return 0;
}
int pihm_nv_ewt_ss(N_Vector y, double rtol, double atol, N_Vector ewt, double *mn, double *ynrm) {
  const long n = L(y);
  RedWork w = X(y)->w;
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_ewt, dim3(nb), dim3(kRedBlock), 0, S(y), n, D(y), rtol, atol, D(ewt), w);
  launch_final(nb, 2, COMB_MIN | (COMB_SUM << 2), w, S(y), nullptr);
  if (pihm_red_wait(w, S(y))) return -2;
  *mn = w.result[0];
  *ynrm = std::sqrt(w.result[1] / (double)n);
  return 0;
}
int pihm_nv_newton_rhs(double rl1, N_Vector zn1, N_Vector acor, double gamma, N_Vector f, N_Vector out) {
  const long n = L(f);
  hipLaunchKernelGGL(k_newton_rhs, dim3(ew_blocks(n)), dim3(kBlock), 0, S(f), n, rl1, D(zn1), D(acor),
                     gamma, D(f), D(out));
  return 0;
}

// ---------------------------------------------------------------- no allocation
// N_VDotProd_Hip on raw device arrays for callers that cannot allocate.  Its
// workspace is a fixed device array, `scratch`: the first pass's partials
// (one per workgroup, at most PIHM_RED_MAXBLOCKS) and, in the last slot, the
// two counters the second pass updates.  The second pass writes the result to
// the caller's device pointer instead of pinned host memory.  Each GPU holds
// its own copy of the array until the process ends; its address on a device
// comes from hipGetSymbolAddress, once per device, with that device current.
}  // extern "C"

namespace {
constexpr int kScratchDoubles = PIHM_RED_MAXBLOCKS + 1;   // partials, then ctr|flag
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::atomic<double *> g_scratch[kMaxDevices];

// Device address of `scratch` on the current device (looked up on the
// device's first call).  The array's name in host code is not a device
// address, so only this pointer is handed to kernels.
int scratch_ptr(double **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -2;
  double *p = g_scratch[dev].load(std::memory_order_acquire);
  if (!p) {
    void *q = nullptr;
    if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q) return -2;
    p = (double *)q;
    g_scratch[dev].store(p, std::memory_order_release);
  }
  *out = p;
  return 0;
}
}  // namespace

extern "C" {

int pihm_nv_dotprod_noalloc(long int n, const double *x_dev, const double *y_dev, double *result_dev,
                            void *hip_stream) {
  if (n < 0 || n > PIHM_NOALLOC_MAX_LENGTH || !result_dev || (n > 0 && (!x_dev || !y_dev))) return -1;
  double *ws = nullptr;
  if (scratch_ptr(&ws)) return -2;
  hipStream_t st = (hipStream_t)hip_stream;
  RedArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = n;
  a.kind = R_DOT;
  a.nvec = 1;
  a.x = x_dev;
  a.ys[0] = y_dev;
  a.work.partials = ws;
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_reduce, dim3(nb), dim3(kRedBlock), 0, st, a);
  // second pass: the same tree as N_VDotProd_Hip's; its counters live in the
  // last slot of the array, the result goes to the caller's device memory
  unsigned *ctr = reinterpret_cast<unsigned *>(ws + PIHM_RED_MAXBLOCKS);
  hipLaunchKernelGGL(k_partials_final, dim3(1), dim3(kRedThreads), 0, st, nb, 1, (unsigned)COMB_SUM,
                     (const double *)ws, result_dev, (double *)nullptr, ctr, (volatile unsigned *)(ctr + 1),
                     (double *)nullptr, 0.0);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int pihm_nv_time_ops(long int n, int iters, double *ms_linearsum, double *ms_dotprod) {
  if (n <= 0 || iters <= 0 || !ms_linearsum || !ms_dotprod) return -1;
  N_Vector x = N_VNew_Hip(n), y = N_VNew_Hip(n), z = N_VNew_Hip(n);
  hipEvent_t ev[2] = {nullptr, nullptr};
  int rc = (x && y && z) ? 0 : -2;
  for (auto &e : ev)
    if (!rc && hipEventCreate(&e) != hipSuccess) rc = -2;
  if (!rc) {
    const hipStream_t st = (hipStream_t)pihm_nv_stream(x);
    N_VConst_Hip(1.0, x);
    N_VConst_Hip(2.0, y);
    N_VLinearSum_Hip(1.5, x, 0.5, y, z);
    double acc = N_VDotProd_Hip(x, y);
    float ms = 0.0f;
    if (hipEventRecord(ev[0], st) != hipSuccess) rc = -2;
    for (int k = 0; k < iters && !rc; k++) N_VLinearSum_Hip(1.5, x, 0.5, y, z);
    if (!rc && (hipEventRecord(ev[1], st) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
                hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess))
      rc = -2;
    if (!rc) *ms_linearsum = ms / iters;
    if (!rc && hipEventRecord(ev[0], st) != hipSuccess) rc = -2;
    for (int k = 0; k < iters && !rc; k++) acc += N_VDotProd_Hip(x, y);
    if (!rc && (hipEventRecord(ev[1], st) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
                hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess))
      rc = -2;
    if (!rc) *ms_dotprod = ms / iters;
    if (!rc && acc != 2.0 * (double)n * (iters + 1)) rc = -2;   // every sum is exact here
  }
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  N_VDestroy_Hip(x);
  N_VDestroy_Hip(y);
  N_VDestroy_Hip(z);
  return rc;
}

}  // extern "C"

namespace {
struct _generic_N_Vector_Ops g_ops = {
    N_VGetVectorID_Hip, N_VClone_Hip,       N_VCloneEmpty_Hip,   N_VDestroy_Hip,
    N_VSpace_Hip,       N_VGetArrayPointer_Hip, N_VSetArrayPointer_Hip, N_VLinearSum_Hip,
    N_VConst_Hip,       N_VProd_Hip,        N_VDiv_Hip,          N_VScale_Hip,
    N_VAbs_Hip,         N_VInv_Hip,         N_VAddConst_Hip,     N_VDotProd_Hip,
    N_VMaxNorm_Hip,     N_VWrmsNorm_Hip,    N_VWrmsNormMask_Hip, N_VMin_Hip,
    N_VWL2Norm_Hip,     N_VL1Norm_Hip,      N_VCompare_Hip,      N_VInvTest_Hip,
    N_VConstrMask_Hip,  N_VMinQuotient_Hip};
struct _generic_N_Vector_Ops *hip_ops() { return &g_ops; }
}  // namespace
