// Small device-side helpers of the step: the stable time step from the CFL
// reduction and the fill of a device scalar, so that a run of steps needs no
// host synchronisation.
#include "wbmhd_common.hpp"

namespace {

// mhd.py:954,963 / mhd3d.py:722-724: dt = cfl * (1 / largest rate); with a
// positive `remaining`, dt = min(dt, remaining) as run() does (mhd.py:1004-1007).
__global__ void k_stable_dt(const double* __restrict__ rate, double cfl, double remaining,
                            double* __restrict__ dt, double* __restrict__ time) {
  if (threadIdx.x == 0) {
    double d = cfl * (1.0 / rate[0]);
    if (remaining > 0.0 && remaining < d) d = remaining;
    dt[0] = d;
    if (time) time[0] = time[0] + d;
  }
}

}  // namespace

extern "C" int wb_stable_dt(const double* rate, double cfl, double remaining, double* dt,
                            double* time, void* stream) {
  hipLaunchKernelGGL(k_stable_dt, dim3(1), dim3(64), 0, (hipStream_t)stream, rate, cfl, remaining,
                     dt, time);
  return wb_last_error();
}
