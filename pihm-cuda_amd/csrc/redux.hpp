// redux.hpp - one-launch element-wise operation fused with NV deterministic sums.
//
// Each workgroup accumulates NV partial sums over a fixed strided set of
// entries (wavefront butterfly, then the four waves in order), writes them,
// and the last workgroup to take a ticket adds the partials in workgroup order
// and writes the NV results to pinned host memory.  The summation tree depends
// only on the length, so every result is reproducible bit for bit.
#pragma once
#include <hip/hip_runtime.h>
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
__global__ __launch_bounds__(kRedThreads) void k_fused_sum(long n, F f, pihm_redwork w) {
  __shared__ double lds[4 * NV];
  __shared__ bool last;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) acc[k] = 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) f(i, acc);
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
    w.partials[(long)k * PIHM_RED_MAXBLOCKS + blockIdx.x] = v;
  }
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = (atomicAdd(w.ticket, 1u) == gridDim.x - 1);
  __syncthreads();
  if (!last) return;
  __threadfence();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double v = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
      v += __builtin_nontemporal_load(&w.partials[(long)k * PIHM_RED_MAXBLOCKS + b]);
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) w.result[k] = ((lds[0] + lds[1]) + lds[2]) + lds[3];
  }
  if (threadIdx.x == 0) *w.ticket = 0u;
}

// Launch and wait; results in out[0..NV).
template <int NV, class F>
int fused_sum(long n, const F &f, pihm_redwork w, hipStream_t st, double *out) {
  hipLaunchKernelGGL((k_fused_sum<NV, F>), dim3(fused_blocks(n)), dim3(kRedThreads), 0, st, n, f, w);
  if (hipGetLastError() != hipSuccess) return -2;
  if (hipStreamSynchronize(st) != hipSuccess) return -2;
  for (int k = 0; k < NV; k++) out[k] = w.result[k];
  return 0;
}

}  // namespace pihm
