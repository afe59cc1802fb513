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

// mhd.py:931-963 with a diffusivity: the smallest of 1 / rate[0] and the explicit
// diffusion limit 0.5 / (D s), D = max(eta, cond / rate[1]) (cond = (gamma - 1)
// kappa_max) and s = dx^-2 + dy^-2 (+ dz^-2) of the cells; with s = 0 the limit is
// 0.5 / rate[2], the largest D s of a hierarchy's patches (wb_amr2d_rate).
__global__ void k_stable_dt_diffusive(const double* __restrict__ rate, double cfl,
                                      double remaining, double eta, double cond, double s,
                                      double* __restrict__ dt, double* __restrict__ time) {
  if (threadIdx.x == 0) {
    double smallest = 1.0 / rate[0];
    double ds;
    if (s > 0.0) {
      const double c = cond / rate[1];
      const double d = c > eta ? c : eta;
      ds = d > 0.0 ? d * s : 0.0;
    } else {
      ds = rate[2];
    }
    if (ds > 0.0) {
      const double limit = 0.5 / ds;
      smallest = limit < smallest ? limit : smallest;
    }
    double d = cfl * smallest;
    if (remaining > 0.0 && remaining < d) d = remaining;
    dt[0] = d;
    if (time) time[0] = time[0] + d;
  }
}

}  // namespace

extern "C" int wb_stable_dt_diffusive(const double* rate, double cfl, double remaining,
                                      double eta, double cond, double s, double* dt, double* time,
                                      void* stream) {
  if (!(eta >= 0.0) || !(cond >= 0.0) || !(s >= 0.0)) return WB_EINVAL;
  hipLaunchKernelGGL(k_stable_dt_diffusive, dim3(1), dim3(64), 0, (hipStream_t)stream, rate, cfl,
                     remaining, eta, cond, s, dt, time);
  return wb_last_error();
}

extern "C" int wb_stable_dt(const double* rate, double cfl, double remaining, double* dt,
                            double* time, void* stream) {
  hipLaunchKernelGGL(k_stable_dt, dim3(1), dim3(64), 0, (hipStream_t)stream, rate, cfl, remaining,
                     dt, time);
  return wb_last_error();
}
