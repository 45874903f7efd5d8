// common.hpp - shared helpers of the MI355X MM-PIHM library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#define PIHM_HIP_TRY(expr)                                                   \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) {                                                  \
      std::fprintf(stderr, "pihm: HIP error %s at %s:%d: %s\n",              \
                   hipGetErrorName(_e), __FILE__, __LINE__, #expr);          \
      return PIHM_ERR_HIP;                                                   \
    }                                                                        \
  } while (0)

namespace pihm {

constexpr int kBlock = 256;      // 4 wavefronts of 64 lanes
constexpr int kNumXcd = 8;       // MI355X: 8 XCDs, workgroups dispatched round-robin

// Maps a hardware workgroup id to a tile id.  Workgroups are dispatched to
// the XCDs round-robin (XCD b % 8 receives b, b+8, b+16, ...).  span 0: each
// XCD works on one contiguous range of tiles, so the Hilbert-ordered
// neighbours of a tile mostly live in the same XCD's L2.  span s >= 1: the
// XCDs take groups of s consecutive tiles in turn (s = 1 is the identity).
// Bijective for any tile count.
__device__ __forceinline__ int xcd_tile(int b, int ntiles, int span = 0) {
  const int x = b % kNumXcd, i = b / kNumXcd;
  if (span <= 0) {
    const int q = ntiles / kNumXcd, r = ntiles % kNumXcd;
    return x * q + (x < r ? x : r) + i;
  }
  const int round = kNumXcd * span, full = ntiles / round;
  if (b >= full * round) return b;           // the last partial round in order
  return ((i / span) * kNumXcd + x) * span + i % span;
}

inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace pihm
