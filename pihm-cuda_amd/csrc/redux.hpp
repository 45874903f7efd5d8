// redux.hpp - element-wise operation fused with NV deterministic sums.
//
// Each workgroup accumulates NV partial sums over a fixed strided set of
// entries (wavefront butterfly, then the four waves in order) and writes
// them; a one-workgroup second launch adds each slot's partials (per thread
// in workgroup order, wavefront butterfly, waves in order) and writes the
// NV results to pinned host memory.  The summation tree
// depends only on the length, so every result is reproducible bit for bit.
// Two launches instead of a last-workgroup ticket: on a multi-XCD device the
// device-scope fence such a ticket needs writes back the whole L2 of the
// issuing XCD, once per workgroup.
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>
#include "common.hpp"
#include "nvector_impl.hpp"

namespace pihm {

constexpr int kRedThreads = 256;

inline int fused_blocks(long n) {
  long b = (n + 2L * kRedThreads - 1) / (2L * kRedThreads);
  if (b < 1) b = 1;
  if (b > PIHM_RED_MAXBLOCKS) b = PIHM_RED_MAXBLOCKS;
  return (int)b;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int NV, class F>
__global__ __launch_bounds__(kRedThreads) void k_fused_sum(long n, F f, double *partials) {
  __shared__ double lds[4 * NV];
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) acc[k] = 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if constexpr (requires(const F &g, long j) { g.load(j); }) {
    // two entries' loads before either entry's stores; the sums keep the
    // order of the entries
    for (; i + stride < n; i += 2 * stride) {
      const auto a = f.load(i);
      const auto b = f.load(i + stride);
      f.apply(i, a, acc);
      f.apply(i + stride, b, acc);
    }
    if (i < n) f.apply(i, f.load(i), acc);
  } else {
    for (; i < n; i += stride) f(i, acc);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const double v = wave_sum(acc[k]);
    if (lane == 0) lds[wave * NV + k] = v;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int k = threadIdx.x;
    const double v = ((lds[k] + lds[NV + k]) + lds[2 * NV + k]) + lds[3 * NV + k];
    partials[(long)k * PIHM_RED_MAXBLOCKS + blockIdx.x] = v;
  }
}

// How the second pass combines the partials of slot k: 2 bits per slot.
enum : unsigned { COMB_SUM = 0u, COMB_MIN = 1u, COMB_MAX = 2u };

__device__ __forceinline__ double comb_ident(unsigned c) {
  return c == COMB_MIN ? DBL_MAX : 0.0;
}
__device__ __forceinline__ double comb(unsigned c, double a, double b) {
  return c == COMB_MIN ? fmin(a, b) : (c == COMB_MAX ? fmax(a, b) : a + b);
}

// Second pass of every reduction: one workgroup adds the nb partials of each
// of the nv slots in workgroup order (strided per thread, wavefront
// butterfly, the four waves in order) and writes the results.  A separate
// launch, so the partials of the first pass are visible without any
// device-wide fence in the first pass.
namespace {
// Every thread loads its partials of all nv slots (workgroups t, t+256,
// t+512, t+768) before combining any, so the second pass waits for one
// round of memory latency; each slot is then combined in the fixed order
// thread (ascending workgroup), wavefront butterfly, waves in order.
// Optionally stores one extra scalar (xstore) for kernels that follow on the
// stream.
__global__ __launch_bounds__(kRedThreads) void k_partials_final(int nb, int nv, unsigned combs,
                                                                 const double *partials, double *result,
                                                                 double *devcopy, unsigned *ctr,
                                                                 volatile unsigned *flag, double *xstore,
                                                                 double xval) {
  constexpr int kPer = PIHM_RED_MAXBLOCKS / kRedThreads;
  static_assert(PIHM_RED_MAXBLOCKS % kRedThreads == 0, "whole rounds of partials per thread");
  __shared__ double lds[(kRedThreads / 64) * PIHM_RED_MAXSUMS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double v[PIHM_RED_MAXSUMS][kPer];
#pragma unroll
  for (int k = 0; k < PIHM_RED_MAXSUMS; k++) {
    if (k < nv) {
      const unsigned c = (combs >> (2 * k)) & 3u;
#pragma unroll
      for (int j = 0; j < kPer; j++) {
        const int b = (int)threadIdx.x + j * kRedThreads;
        v[k][j] = (b < nb) ? partials[(long)k * PIHM_RED_MAXBLOCKS + b] : comb_ident(c);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < PIHM_RED_MAXSUMS; k++) {
    if (k < nv) {
      const unsigned c = (combs >> (2 * k)) & 3u;
      double x = v[k][0];
#pragma unroll
      for (int j = 1; j < kPer; j++) x = comb(c, x, v[k][j]);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) x = comb(c, x, __shfl_xor(x, off, 64));
      if (lane == 0) lds[wave * PIHM_RED_MAXSUMS + k] = x;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < nv) {
    const int k = threadIdx.x;
    const unsigned c = (combs >> (2 * k)) & 3u;
    double r = lds[k];
#pragma unroll
    for (int w2 = 1; w2 < kRedThreads / 64; w2++) r = comb(c, r, lds[w2 * PIHM_RED_MAXSUMS + k]);
    result[k] = r;
    if (devcopy) devcopy[k] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (xstore) *xstore = xval;
    // results reach host memory before the count that announces them
    __threadfence_system();
    const unsigned s = *ctr + 1u;
    *ctr = s;
    *flag = s;   // made visible by the end of the kernel
  }
}
}  // namespace

// Launch the second pass of a reduction whose first pass wrote nb partials.
inline void launch_final(int nb, int nv, unsigned combs, pihm_redwork w, hipStream_t st, double *devcopy,
                         double *xstore = nullptr, double xval = 0.0) {
  hipLaunchKernelGGL(k_partials_final, dim3(1), dim3(kRedThreads), 0, st, nb, nv, combs, w.partials, w.result,
                     devcopy, w.ctr, w.flag, xstore, xval);
}

// Launch both passes without waiting (devcopy: optional device copy of the
// results; xstore: optional scalar store made by the second pass).
template <int NV, class F>
void fused_sum_launch(long n, const F &f, pihm_redwork w, hipStream_t st, double *devcopy,
                      double *xstore = nullptr, double xval = 0.0) {
  const int nb = fused_blocks(n);
  hipLaunchKernelGGL((k_fused_sum<NV, F>), dim3(nb), dim3(kRedThreads), 0, st, n, f, w.partials);
  launch_final(nb, NV, 0u, w, st, devcopy, xstore, xval);
}

// Launch both passes and wait; results in out[0..NV).
template <int NV, class F>
int fused_sum(long n, const F &f, pihm_redwork w, hipStream_t st, double *out, double *devcopy = nullptr,
              double *xstore = nullptr, double xval = 0.0) {
  fused_sum_launch<NV>(n, f, w, st, devcopy, xstore, xval);
  if (hipGetLastError() != hipSuccess) return -2;
  if (pihm_red_wait(w, st)) return -2;
  for (int k = 0; k < NV; k++) out[k] = w.result[k];
  return 0;
}

}  // namespace pihm
