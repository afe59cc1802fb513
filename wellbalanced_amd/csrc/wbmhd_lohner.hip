// The second-derivative refinement estimator (Löhner 1987, in the normalised form
// of FLASH, Fryxell et al. 2000 sec. 6) on a haloed 3D block: the label source of
// the learned refinement indicator and its fallback.  FP64, -ffp-contract=off.
//
// For u in {rho, p} at an interior cell, with d_a(j) = u(j+e_a) - u(j-e_a) and
// s_a(j) = |u(j+e_a)| + |u(j-e_a)|:
//   num = sum_ab (d_a(i+e_b) - d_a(i-e_b))^2
//   den = sum_ab (|d_a(i+e_b)| + |d_a(i-e_b)| + eps (s_a(i+e_b) + s_a(i-e_b)))^2
//   E = sqrt(num / den) (0 where den = 0); the estimate is max(E(rho), E(p)).
#include "mhd_device.hpp"
#include "wbmhd_common.hpp"

using namespace wb;

namespace {

struct G3 {
  int X, Y, Z;
};

__device__ __forceinline__ double AXf(const double* ax, const G3& g, int i, int j, int k) {
  return ax[((size_t)i * (g.Y + 1) + j) * (g.Z + 1) + k];
}
__device__ __forceinline__ double AYf(const double* ay, const G3& g, int i, int j, int k) {
  return ay[((size_t)i * g.Y + j) * (g.Z + 1) + k];
}
__device__ __forceinline__ double AZf(const double* az, const G3& g, int i, int j, int k) {
  return az[((size_t)i * (g.Y + 1) + j) * g.Z + k];
}

// rho and p of every cell of the haloed block (p as in mhd3d.py:475-495).
__global__ void k_rho_p3d(const double* __restrict__ state, G3 g, double dx, double dy, double dz,
                          double b0x, double b0y, double b0z, double gamma,
                          double* __restrict__ out) {
  const size_t P = (size_t)g.X * g.Y * g.Z;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= P) return;
  const int k = id % g.Z, t = id / g.Z, j = t % g.Y, i = t / g.Y;
  const double* ax = state + 5 * P;
  const double* ay = ax + (size_t)g.X * (g.Y + 1) * (g.Z + 1);
  const double* az = ay + (size_t)(g.X + 1) * g.Y * (g.Z + 1);
  auto bxf = [&](int ii) {
    return (AZf(az, g, ii, j + 1, k) - AZf(az, g, ii, j, k)) / dy - (AYf(ay, g, ii, j, k + 1) - AYf(ay, g, ii, j, k)) / dz + b0x;
  };
  auto byf = [&](int jj) {
    return (AXf(ax, g, i, jj, k + 1) - AXf(ax, g, i, jj, k)) / dz - (AZf(az, g, i + 1, jj, k) - AZf(az, g, i, jj, k)) / dx + b0y;
  };
  auto bzf = [&](int kk) {
    return (AYf(ay, g, i + 1, j, kk) - AYf(ay, g, i, j, kk)) / dx - (AXf(ax, g, i, j + 1, kk) - AXf(ax, g, i, j, kk)) / dy + b0z;
  };
  const double bx = 0.5 * (bxf(i) + bxf(i + 1));
  const double by = 0.5 * (byf(j) + byf(j + 1));
  const double bz = 0.5 * (bzf(k) + bzf(k + 1));
  const double r = state[id];
  const double u = state[P + id] / r, v = state[2 * P + id] / r, w = state[3 * P + id] / r;
  const double p = (gamma - 1.0) * (state[4 * P + id] - 0.5 * r * (u * u + v * v + w * w) -
                                    0.5 * (bx * bx + by * by + bz * bz));
  out[id] = r;
  out[P + id] = p;
}

__device__ __forceinline__ double lohner_at(const double* __restrict__ u, const G3& g, int i, int j,
                                            int k, double eps) {
  const size_t st[3] = {(size_t)g.Y * g.Z, (size_t)g.Z, 1};
  const size_t c = ((size_t)i * g.Y + j) * g.Z + k;
  double num = 0.0, den = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const size_t hi = c + st[b], lo = c - st[b];
      const double dhi = u[hi + st[a]] - u[hi - st[a]];
      const double dlo = u[lo + st[a]] - u[lo - st[a]];
      const double shi = __builtin_fabs(u[hi + st[a]]) + __builtin_fabs(u[hi - st[a]]);
      const double slo = __builtin_fabs(u[lo + st[a]]) + __builtin_fabs(u[lo - st[a]]);
      num = num + (dhi - dlo) * (dhi - dlo);
      const double t = __builtin_fabs(dhi) + __builtin_fabs(dlo) + eps * (shi + slo);
      den = den + t * t;
    }
  }
  return den > 0.0 ? __builtin_sqrt(num / den) : 0.0;
}

// The estimate of every interior cell, out (nx, ny, nz).
__global__ void k_lohner3d(const double* __restrict__ rp, G3 g, double eps,
                           double* __restrict__ out) {
  const int nx = g.X - 4, ny = g.Y - 4, nz = g.Z - 4;
  const size_t n = (size_t)nx * ny * nz;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const int c = id % nz, t = id / nz, b = t % ny, a = t / ny;
  const size_t P = (size_t)g.X * g.Y * g.Z;
  const double er = lohner_at(rp, g, a + 2, b + 2, c + 2, eps);
  const double ep = lohner_at(rp + P, g, a + 2, b + 2, c + 2, eps);
  out[id] = er > ep ? er : ep;
}

}  // namespace

extern "C" int wb_lohner3d(const double* state, int64_t nx, int64_t ny, int64_t nz, double dx,
                           double dy, double dz, double b0x, double b0y, double b0z, double gamma,
                           double eps, double* scratch, double* out, void* stream) {
  if (nx < 1 || ny < 1 || nz < 1) return WB_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  G3 g{(int)nx + 4, (int)ny + 4, (int)nz + 4};
  const size_t P = (size_t)g.X * g.Y * g.Z;
  hipLaunchKernelGGL(k_rho_p3d, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, state, g, dx,
                     dy, dz, b0x, b0y, b0z, gamma, scratch);
  const size_t n = (size_t)nx * ny * nz;
  hipLaunchKernelGGL(k_lohner3d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, scratch, g,
                     eps, out);
  return wb_last_error();
}
