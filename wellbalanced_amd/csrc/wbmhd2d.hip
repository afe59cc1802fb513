// 2D ideal-MHD kernels for gfx950: the reference's four whole-block kernels as
// device kernels with the same arguments (per-kernel twins), plus the halo fill,
// the wall fix-up, the conservative/CT update with the fused Heun average and the
// CFL reduction.  FP64 throughout; built with -ffp-contract=off.
//
// Layout: C order, last index contiguous (y), one plane per variable.  A haloed
// block is X = nx + 4 by Y = ny + 4 cells; A_z sits on its (X+1) x (Y+1) nodes.
// The model state is one buffer: 6 cell planes (rho, mx, my, mz, E, bz) then A_z.
#include "mhd_device.hpp"
#include "wbmhd_common.hpp"

using namespace wb;

namespace {

// mhd_primitive_pyccel.py:23-87.  One thread per node-grid point (i, j) in
// (X+1) x (Y+1): the x-face field if j < Y, the y-face field if i < X, the cell
// if both.
__global__ void k_primitive2d(const double* __restrict__ rho, const double* __restrict__ mx,
                              const double* __restrict__ my, const double* __restrict__ mz,
                              const double* __restrict__ en, const double* __restrict__ bz,
                              const double* __restrict__ az, int X, int Y, double dx, double dy,
                              double b0x, double b0y, double gamma, double* __restrict__ prim,
                              double* __restrict__ bxf, double* __restrict__ byf,
                              double* __restrict__ ecc, int* __restrict__ bad) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (i > X || j > Y) return;
  const int YN = Y + 1;
  if (j < Y) bxf[(size_t)i * Y + j] = (az[(size_t)i * YN + j + 1] - az[(size_t)i * YN + j]) / dy + b0x;
  if (i < X) byf[(size_t)i * YN + j] = -(az[(size_t)(i + 1) * YN + j] - az[(size_t)i * YN + j]) / dx + b0y;
  if (i >= X || j >= Y) return;
  const size_t c = (size_t)i * Y + j;
  const double bx0 = (az[(size_t)i * YN + j + 1] - az[(size_t)i * YN + j]) / dy + b0x;
  const double bx1 = (az[(size_t)(i + 1) * YN + j + 1] - az[(size_t)(i + 1) * YN + j]) / dy + b0x;
  const double by0 = -(az[(size_t)(i + 1) * YN + j] - az[(size_t)i * YN + j]) / dx + b0y;
  const double by1 = -(az[(size_t)(i + 1) * YN + j + 1] - az[(size_t)i * YN + j + 1]) / dx + b0y;
  const double r = rho[c];
  const double u = mx[c] / r, v = my[c] / r, w = mz[c] / r;
  const double bx = 0.5 * (bx0 + bx1);
  const double by = 0.5 * (by0 + by1);
  const double b = bz[c];
  const double p = (gamma - 1.0) * (en[c] - 0.5 * r * (u * u + v * v + w * w) -
                                    0.5 * (bx * bx + by * by + b * b));
  const size_t P = (size_t)X * Y;
  prim[c] = r;
  prim[P + c] = u;
  prim[2 * P + c] = v;
  prim[3 * P + c] = w;
  prim[4 * P + c] = p;
  prim[5 * P + c] = bx;
  prim[6 * P + c] = by;
  prim[7 * P + c] = b;
  ecc[c] = bx * v - by * u;
  bool is_bad = !(r > 0.0 && p > 0.0) ||
                !(finite(r) && finite(u) && finite(v) && finite(w) && finite(p) && finite(bx) &&
                  finite(by) && finite(b));
  if (is_bad) atomicAdd(bad, 1);
}

// mhd_sweep_pyccel.py:254-319.  One thread per face (k, line); axis 0 writes
// flux (7, X-3, Y), axis 1 flux (7, X, Y-3).
template <int AXIS>
__global__ void __launch_bounds__(128) k_sweep2d(const double* __restrict__ prim, const double* __restrict__ bn, int X,
                          int Y, int limiter, double gamma, double* __restrict__ flux) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y;
  const size_t P = (size_t)X * Y;
  // normal-frame rows: rho, v_n, v_t1, v_t2, p, B_t1, B_t2 (mhd_sweep_pyccel.py:268-275)
  constexpr int comp[7] = {0, 1 + AXIS, 2 - AXIS, 3, 4, 6 - AXIS, 7};
  double q[7][4];
  double normal;
  size_t out, ostride;
  if (AXIS == 0) {
    const int k = row, line = t;  // faces along x, lines along y (contiguous)
    if (k >= X - 3 || line >= Y) return;
#pragma unroll
    for (int c = 0; c < 7; ++c)
#pragma unroll
      for (int s = 0; s < 4; ++s) q[c][s] = prim[comp[c] * P + (size_t)(k + s) * Y + line];
    normal = bn[(size_t)(k + 2) * Y + line];
    out = (size_t)k * Y + line;
    ostride = (size_t)(X - 3) * Y;
  } else {
    const int line = row, k = t;
    if (k >= Y - 3 || line >= X) return;
#pragma unroll
    for (int c = 0; c < 7; ++c)
#pragma unroll
      for (int s = 0; s < 4; ++s) q[c][s] = prim[comp[c] * P + (size_t)line * Y + k + s];
    normal = bn[(size_t)line * (Y + 1) + k + 2];
    out = (size_t)line * (Y - 3) + k;
    ostride = (size_t)X * (Y - 3);
  }
  double L[7], R[7], f[7];
  reconstruct<double>(q, limiter, L, R);
  hlld<double>(L, R, normal, gamma, f);
#pragma unroll
  for (int c = 0; c < 7; ++c) flux[c * ostride + out] = f[c];
}

// ct_contact_pyccel.py:26-61.  One thread per node (a, b) of the block.
__global__ void k_ct2d(const double* __restrict__ fx_by, const double* __restrict__ gy_bx,
                       const double* __restrict__ fx_rho, const double* __restrict__ gy_rho,
                       const double* __restrict__ ecc, int X, int Y, double* __restrict__ ez) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  const int NA = X - 3, NB = Y - 3;
  if (a >= NA || b >= NB) return;
  const int i = a + 2, j = b + 2;
  const int YG = Y - 3;  // gy row length
  ez[(size_t)a * NB + b] = ct_node(
      fx_by[(size_t)(i - 2) * Y + j - 1], fx_by[(size_t)(i - 2) * Y + j],
      gy_bx[(size_t)(i - 1) * YG + j - 2], gy_bx[(size_t)i * YG + j - 2],
      fx_rho[(size_t)(i - 2) * Y + j - 1], fx_rho[(size_t)(i - 2) * Y + j],
      gy_rho[(size_t)(i - 1) * YG + j - 2], gy_rho[(size_t)i * YG + j - 2],
      ecc[(size_t)(i - 1) * Y + j - 1], ecc[(size_t)i * Y + j - 1], ecc[(size_t)(i - 1) * Y + j],
      ecc[(size_t)i * Y + j]);
}

// mhd_rate_pyccel.py:30-83: the largest fast-wave rate and the smallest density
// of the block's own cells, reduced into ordered-integer keys (order-independent,
// so exact).
__global__ void __launch_bounds__(256) k_rate2d(const double* __restrict__ rho, const double* __restrict__ mx,
                         const double* __restrict__ my, const double* __restrict__ mz,
                         const double* __restrict__ en, const double* __restrict__ bz,
                         const double* __restrict__ az, int X, int Y, int h, double dx,
                         double dy, double b0x, double b0y, double gamma,
                         unsigned long long* __restrict__ keys) {
  const int nx = X - 2 * h, ny = Y - 2 * h;
  const unsigned total = (unsigned)nx * ny;
  double rate = 0.0, rmin = __builtin_inf();
  for (unsigned id = blockIdx.x * blockDim.x + threadIdx.x; id < total; id += gridDim.x * blockDim.x) {
    const int b = id % ny, a = id / ny;
    const int i = a + h, j = b + h, YN = Y + 1;
    const size_t c = (size_t)i * Y + j;
    const double r = rho[c];
    const double u = mx[c] / r, v = my[c] / r, w = mz[c] / r;
    const double bx = 0.5 * ((az[(size_t)i * YN + j + 1] - az[(size_t)i * YN + j]) / dy + b0x +
                             ((az[(size_t)(i + 1) * YN + j + 1] - az[(size_t)(i + 1) * YN + j]) / dy + b0x));
    const double by = 0.5 * (-(az[(size_t)(i + 1) * YN + j] - az[(size_t)i * YN + j]) / dx + b0y +
                             (-(az[(size_t)(i + 1) * YN + j + 1] - az[(size_t)i * YN + j + 1]) / dx + b0y));
    const double bzc = bz[c];
    const double p = (gamma - 1.0) * (en[c] - 0.5 * r * (u * u + v * v + w * w) -
                                      0.5 * (bx * bx + by * by + bzc * bzc));
    const double rc = (__builtin_fabs(u) + fast_speed_rate(r, p, bx, by, bzc, gamma)) / dx +
                      (__builtin_fabs(v) + fast_speed_rate(r, p, by, bx, bzc, gamma)) / dy;
    if (rc > rate) rate = rc;
    if (r < rmin) rmin = r;
  }
  block_reduce_rate_rho_once(rate, rmin, keys);
}

// Halo pass along one axis for the 6 cell planes and the A_z node plane.
// kind: 0 periodic, 1 outflow, 2 wall.  Threads: ghost slot g in [0, 5) x the
// full extent of the other axis x the 7 arrays.
__global__ void k_halo2d(double* __restrict__ state, int nx, int ny, int axis, int kind) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y;          // 0,1: low ghosts; 2,3,4: high slots
  const int field = blockIdx.z;      // 0..5 cells, 6 = A_z
  const int X = nx + 4, Y = ny + 4;
  const bool node = field == 6;
  const int len[2] = {node ? X + 1 : X, node ? Y + 1 : Y};
  const int other = len[1 - axis];
  if (t >= other) return;
  double* a = state + (size_t)field * X * Y;
  const int L = len[axis];
  const size_t sa = axis == 0 ? (size_t)len[1] : 1;     // stride along axis
  const size_t so = axis == 0 ? 1 : (size_t)len[1];     // stride along the other axis
  double* line = a + t * so;
  const int n = axis == 0 ? nx : ny;
  const int h = HALO;
  auto at = [&](int idx) -> double& { return line[(size_t)idx * sa]; };
  if (kind == 0) {
    if (g < 2) {
      at(g) = at(g + n);
    } else {
      const int tt = g - 2;
      if (!node && tt >= 2) return;
      at(n + h + tt) = at(h + tt);
    }
  } else if (kind == 1) {
    if (g >= 4) return;
    if (!node) {
      if (g < 2) at(g) = at(h);
      else at(L - 4 + g) = at(L - h - 1);   // L-2, L-1
    } else {
      if (g < 2) {
        const int k = h - g;  // slot 0 -> k=2, slot 1 -> k=1
        const double edge = at(h), inner = at(h + 1);
        at(h - k) = edge + (double)k * (edge - inner);
      } else {
        const int k = g - 1;  // slot 2 -> k=1, slot 3 -> k=2
        const double edge = at(L - h - 1), inner = at(L - h - 2);
        at(L - h - 1 + k) = edge + (double)k * (edge - inner);
      }
    }
  } else {
    if (g >= 4) return;
    if (!node) {
      const double sign = (field == 1 + axis) ? -1.0 : 1.0;
      if (g < 2) {
        const int k = g;  // m[h-1-k] = sign*m[h+k]
        at(h - 1 - k) = sign * at(h + k);
      } else {
        const int k = g - 2;  // m[L-h+k] = sign*m[L-h-1-k]
        at(L - h + k) = sign * at(L - h - 1 - k);
      }
    } else {
      if (g < 2) {
        const int k = g + 1;
        const double wall = at(h);
        at(h - k) = 2.0 * wall - at(h + k);
      } else {
        const int k = g - 1;
        const double wall = at(L - h - 1);
        at(L - h - 1 + k) = 2.0 * wall - at(L - h - 1 - k);
      }
    }
  }
}

// Nothing but the pressure crosses a wall and E_z along it is zero
// (mhd.py:873-886): rows 0, 2, 3, 4, 6 of the wall-face fluxes and the wall
// row of E_z are zeroed.  mask bits: 1 x-lower, 2 x-upper, 4 y-lower, 8 y-upper.
__global__ void k_walls2d(double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ ez,
                          int nx, int ny, int mask) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int X = nx + 4, Y = ny + 4;
  constexpr int rows[5] = {0, 2, 3, 4, 6};
  if (mask & 3) {
    if (t < Y) {
      const size_t s = (size_t)(X - 3) * Y;
      for (int r = 0; r < 5; ++r) {
        if (mask & 1) fx[rows[r] * s + t] = 0.0;
        if (mask & 2) fx[rows[r] * s + (size_t)nx * Y + t] = 0.0;
      }
    }
    if (t <= ny) {
      if (mask & 1) ez[t] = 0.0;
      if (mask & 2) ez[(size_t)nx * (ny + 1) + t] = 0.0;
    }
  }
  if (mask & 12) {
    if (t < X) {
      const size_t s = (size_t)X * (Y - 3);
      for (int r = 0; r < 5; ++r) {
        if (mask & 4) fy[rows[r] * s + (size_t)t * (Y - 3)] = 0.0;
        if (mask & 8) fy[rows[r] * s + (size_t)t * (Y - 3) + ny] = 0.0;
      }
    }
    if (t <= nx) {
      if (mask & 4) ez[(size_t)t * (ny + 1)] = 0.0;
      if (mask & 8) ez[(size_t)t * (ny + 1) + ny] = 0.0;
    }
  }
}

// The stage update u' = u - (dt/dx (Fx+ - Fx-) + dt/dy (Fy+ - Fy-)) of the 6
// cell fields and A_z' = A_z - dt E_z, from state `in` into `out`.  With heun,
// out = (u' + out) * 0.5 (mhd.py:983-988) reading out's own previous value.
__global__ void k_update2d(const double* __restrict__ in, double* __restrict__ out,
                           const double* __restrict__ fx, const double* __restrict__ fy,
                           const double* __restrict__ ez, int nx, int ny, double dx, double dy,
                           const double* __restrict__ dtp, int heun) {
  const double dt = *dtp;
  const double dtdx = dt / dx, dtdy = dt / dy;
  constexpr int XR[6] = {0, 1, 2, 3, 4, 6};
  constexpr int YR[6] = {0, 2, 1, 3, 4, 6};
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  const int X = nx + 4, Y = ny + 4, h = HALO;
  const size_t P = (size_t)X * Y;
  if (a < nx && b < ny) {
    const size_t sx = (size_t)(X - 3) * Y, sy = (size_t)X * (Y - 3);
    const size_t c = (size_t)(a + h) * Y + b + h;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const double* gx = fx + XR[f] * sx;
      const double* gy = fy + YR[f] * sy;
      const double dxf = gx[(size_t)(a + 1) * Y + b + h] - gx[(size_t)a * Y + b + h];
      const double dyf = gy[(size_t)(a + h) * (Y - 3) + b + 1] - gy[(size_t)(a + h) * (Y - 3) + b];
      double v = in[f * P + c] - (dtdx * dxf + dtdy * dyf);
      if (heun) v = (v + out[f * P + c]) * 0.5;
      out[f * P + c] = v;
    }
  }
  if (a <= nx && b <= ny) {
    const size_t n = 6 * P + (size_t)(a + h) * (Y + 1) + b + h;
    double v = in[n] - dt * ez[(size_t)a * (ny + 1) + b];
    if (heun) v = (v + out[n]) * 0.5;
    out[n] = v;
  }
}

}  // namespace

extern "C" {

int wb_primitive2d(const double* rho, const double* mx, const double* my, const double* mz,
                   const double* energy, const double* bz, const double* az, int64_t X, int64_t Y,
                   double dx, double dy, double b0x, double b0y, double gamma, double* prim,
                   double* bxf, double* byf, double* ecc, int32_t* bad, void* stream) {
  if (X < 1 || Y < 1 || X > INT32_MAX / 2 || Y > INT32_MAX / 2) return WB_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, bad, 1);
  dim3 grid((unsigned)((Y + 1 + 255) / 256), (unsigned)(X + 1));
  hipLaunchKernelGGL(k_primitive2d, grid, dim3(256), 0, s, rho, mx, my, mz, energy, bz, az, (int)X,
                     (int)Y, dx, dy, b0x, b0y, gamma, prim, bxf, byf, ecc, (int*)bad);
  return wb_last_error();
}

int wb_sweep2d(const double* prim, const double* bn, int64_t X, int64_t Y, int axis, int limiter,
               double gamma, double* flux, void* stream) {
  if (X < 4 || Y < 4 || (axis != 0 && axis != 1) || limiter < 0 || limiter > 3) return WB_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  if (axis == 0) {
    dim3 grid((unsigned)((Y + 127) / 128), (unsigned)(X - 3));
    hipLaunchKernelGGL(k_sweep2d<0>, grid, dim3(128), 0, s, prim, bn, (int)X, (int)Y, limiter,
                       gamma, flux);
  } else {
    dim3 grid((unsigned)((Y - 3 + 127) / 128), (unsigned)X);
    hipLaunchKernelGGL(k_sweep2d<1>, grid, dim3(128), 0, s, prim, bn, (int)X, (int)Y, limiter,
                       gamma, flux);
  }
  return wb_last_error();
}

int wb_ct_contact(const double* fx_by, const double* gy_bx, const double* fx_rho,
                  const double* gy_rho, const double* ecc, int64_t X, int64_t Y, double* ez,
                  void* stream) {
  if (X < 4 || Y < 4) return WB_EINVAL;
  dim3 grid((unsigned)((Y - 3 + 255) / 256), (unsigned)(X - 3));
  hipLaunchKernelGGL(k_ct2d, grid, dim3(256), 0, (hipStream_t)stream, fx_by, gy_bx, fx_rho, gy_rho,
                     ecc, (int)X, (int)Y, ez);
  return wb_last_error();
}

int wb_rate2d(const double* rho, const double* mx, const double* my, const double* mz,
              const double* energy, const double* bz, const double* az, int64_t X, int64_t Y,
              int halo, double dx, double dy, double b0x, double b0y, double gamma,
              double* out, void* stream) {
  if (halo < 0 || X - 2 * halo < 1 || Y - 2 * halo < 1) return WB_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  unsigned long long* keys = (unsigned long long*)out;
  hipLaunchKernelGGL(k_rate_init, dim3(1), dim3(64), 0, s, keys);
  const int64_t total = (X - 2 * halo) * (Y - 2 * halo);
  if (total >= ((int64_t)1 << 31)) return WB_EINVAL;
  const int64_t want = (total + 255) / 256;
  dim3 grid((unsigned)(want < 2048 ? want : 2048));
  hipLaunchKernelGGL(k_rate2d, grid, dim3(256), 0, s, rho, mx, my, mz, energy, bz, az, (int)X,
                     (int)Y, halo, dx, dy, b0x, b0y, gamma, keys);
  hipLaunchKernelGGL(k_rate_finish, dim3(1), dim3(64), 0, s, keys, out);
  return wb_last_error();
}

int wb_mhd2d_fill_halos(double* state, int64_t nx, int64_t ny, int periodic_mask, int wall_mask,
                        void* stream) {
  if (nx < 2 || ny < 2) return WB_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const int64_t n[2] = {nx, ny};
  for (int pass = 0; pass < 2; ++pass) {
    for (int axis = 0; axis < 2; ++axis) {
      const bool periodic = (periodic_mask >> axis) & 1;
      if ((pass == 0) != periodic) continue;
      const int kind = periodic ? 0 : (((wall_mask >> axis) & 1) ? 2 : 1);
      const int64_t other = n[1 - axis] + 5;
      dim3 grid((unsigned)((other + 255) / 256), 5, 7);
      hipLaunchKernelGGL(k_halo2d, grid, dim3(256), 0, s, state, (int)nx, (int)ny, axis, kind);
    }
  }
  return wb_last_error();
}

int wb_mhd2d_walls(double* fx, double* fy, double* ez, int64_t nx, int64_t ny, int side_mask,
                   void* stream) {
  if (side_mask == 0) return 0;
  const int64_t m = (nx > ny ? nx : ny) + 5;
  hipLaunchKernelGGL(k_walls2d, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, fx, fy, ez, (int)nx, (int)ny, side_mask);
  return wb_last_error();
}

int wb_mhd2d_update(const double* in, double* out, const double* fx, const double* fy,
                    const double* ez, int64_t nx, int64_t ny, const double* dt, double dx,
                    double dy, int heun, void* stream) {
  if (nx < 1 || ny < 1) return WB_EINVAL;
  dim3 grid((unsigned)((ny + 1 + 255) / 256), (unsigned)(nx + 1));
  hipLaunchKernelGGL(k_update2d, grid, dim3(256), 0, (hipStream_t)stream, in, out, fx, fy, ez,
                     (int)nx, (int)ny, dx, dy, dt, heun);
  return wb_last_error();
}

}  // extern "C"
