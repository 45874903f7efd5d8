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

// Maps a hardware workgroup id to a tile id so that each XCD (which receives
// workgroups b, b+8, b+16, ...) works on one contiguous range of tiles; the
// Hilbert-ordered neighbours of a tile then mostly live in the same XCD's L2.
// Bijective for any tile count.
__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
  int q = ntiles / kNumXcd, r = ntiles % kNumXcd;
  int x = b % kNumXcd, i = b / kNumXcd;
  return x * q + (x < r ? x : r) + i;
}

inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace pihm
