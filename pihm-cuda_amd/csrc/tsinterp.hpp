// tsinterp.hpp - linear interpolation of a forcing time series on the device.
//
// IntrplForc (src/forcing.c:482-530): bisection over records 1..nrec-1 for the
// interval [ftime[m-1], ftime[m]] that holds t, then
//   ((ftime[m] - t) * d[m-1] + (t - ftime[m-1]) * d[m]) / (ftime[m] - ftime[m-1])
// per variable.  The same bisection picks the same interval for a time equal
// to a record time, so the result rounds as the reference's does.  The caller
// checks ftime[0] <= t <= ftime[nrec-1] on the host.
#pragma once

#include <hip/hip_runtime.h>

namespace pihm {

// One series: ft[nrec] ascending, d[nrec * nvar] record-major; out[v * ostride].
template <int NV>
__device__ inline void ts_interp(const double *ft, const double *d, int nrec, double t, double *out,
                                 long ostride) {
  int first = 1, last = nrec - 1;
  while (first <= last) {
    const int middle = (first + last) / 2;
    if (t >= ft[middle - 1] && t <= ft[middle]) {
#pragma unroll
      for (int j = 0; j < NV; j++)
        out[j * ostride] = ((ft[middle] - t) * d[(long)(middle - 1) * NV + j] + (t - ft[middle - 1]) * d[(long)middle * NV + j]) /
                           (ft[middle] - ft[middle - 1]);
      return;
    } else if (ft[middle] > t) {
      last = middle - 1;
    } else {
      first = middle + 1;
    }
  }
}

}  // namespace pihm
