// The CFL rate of a uniform 3D grid (mhd3d.py:674-724): the largest
// sum_a (|u_a| + c_f,a) / dx_a over the interior cells and the smallest density,
// reduced exactly (order-independent maxima of ordered keys, one atomic pair per
// workgroup).
//
// A thread owns one (y, z) column and marches along x over RATE_SEG planes: the
// A_y and A_z values of plane i+1 it loads are those of plane i in the next cell,
// so a cell costs 13 loads (4 A_x, 2 A_y, 2 A_z, 5 conserved) with no index
// division; a wave covers 64 consecutive z (one 512-byte run per load).  FAST (the
// mixed path) divides once per cell and multiplies by reciprocals; otherwise every
// division of the reference is kept, so the rate is bit-for-bit the reference's.
#pragma once

#include "stage3d.hpp"

namespace wb {

constexpr int RATE_SEG = 16;

// Column (j, k) of the state `a` over planes i0 .. i1 - 1 (haloed indices): the largest
// rate into `rate`, the smallest density into `rmin`.
template <bool FAST>
__device__ __forceinline__ void rate_column(const StageArgs& a, int j, int k, int i0, int i1,
                                            double& rate, double& rmin) {
  {
    const Addr3 d = addr3(a);
    const char* in = reinterpret_cast<const char*>(a.in);
    const uint32_t c0 = (uint32_t)j * d.crow + (uint32_t)k * 8u;  // cells, A_z
    const uint32_t c1 = (uint32_t)j * d.arow + (uint32_t)k * 8u;  // A_x, A_y
    const double rdx = 1.0 / a.dx, rdy = 1.0 / a.dy, rdz = 1.0 / a.dz;
    // A_y(i, j, k + t) and A_z(i, j + t, k) of the current plane
    const char* py = in + d.oy + (size_t)i0 * d.yplane;
    const char* pz = in + d.oz + (size_t)i0 * d.zplane;
    double ay0 = ldb(py, c1), ay1 = ldb(py + 8, c1);
    double az0 = ldb(pz, c0), az1 = ldb(pz + d.crow, c0);
    for (int i = i0; i < i1; ++i) {
      const char* px = in + d.ox + (size_t)i * d.xplane;
      const char* qy = in + d.oy + (size_t)(i + 1) * d.yplane;
      const char* qz = in + d.oz + (size_t)(i + 1) * d.zplane;
      const char* pc = in + (size_t)i * d.cplane;
      const double ax[2][2] = {{ldb(px, c1), ldb(px + 8, c1)},
                               {ldb(px + d.arow, c1), ldb(px + d.arow + 8, c1)}};
      const double ay[2][2] = {{ay0, ay1}, {ldb(qy, c1), ldb(qy + 8, c1)}};
      const double az[2][2] = {{az0, az1}, {ldb(qz, c0), ldb(qz + d.crow, c0)}};
      double cv[5];
#pragma unroll
      for (int q = 0; q < 5; ++q) cv[q] = ldb(pc + q * d.field, c0);
      CellPrim<FAST> c;
      c.finish(a, ax, ay, az, cv, rdx, rdy, rdz);
      double rc;
      if constexpr (FAST) {
        const double ri = 1.0 / c.r;
        const double a2 = a.gamma * c.p * ri;
        const double s = a2 + (c.bx * c.bx + c.by * c.by + c.bz * c.bz) * ri;
        const double q4 = 4.0 * a2 * ri;
        auto cf = [&](double bn) {
          const double disc = s * s - q4 * (bn * bn);
          return __builtin_sqrt(0.5 * (s + __builtin_sqrt(__builtin_fmax(disc, 0.0))));
        };
        rc = (__builtin_fabs(c.u) + cf(c.bx)) * rdx + (__builtin_fabs(c.v) + cf(c.by)) * rdy +
             (__builtin_fabs(c.w) + cf(c.bz)) * rdz;
      } else {
        rc = (__builtin_fabs(c.u) + fast_speed_rate(c.r, c.p, c.bx, c.by, c.bz, a.gamma)) / a.dx +
             (__builtin_fabs(c.v) + fast_speed_rate(c.r, c.p, c.by, c.bz, c.bx, a.gamma)) / a.dy +
             (__builtin_fabs(c.w) + fast_speed_rate(c.r, c.p, c.bz, c.bx, c.by, a.gamma)) / a.dz;
      }
      if (rc > rate) rate = rc;
      if (c.r < rmin) rmin = c.r;
      ay0 = ay[1][0];
      ay1 = ay[1][1];
      az0 = az[1][0];
      az1 = az[1][1];
    }
  }
}

template <bool FAST>
__global__ void __launch_bounds__(256) k_rate3d_columns(StageArgs a,
                                                       unsigned long long* __restrict__ keys) {
  const Grid3 g = a.g;
  const int k = 2 + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = 2 + blockIdx.y * 4 + (threadIdx.x >> 6);
  const int i0 = 2 + blockIdx.z * RATE_SEG;
  const int i1 = min(i0 + RATE_SEG, g.nx + 2);
  double rate = 0.0, rmin = __builtin_inf();
  if (j < g.ny + 2 && k < g.nz + 2) rate_column<FAST>(a, j, k, i0, i1, rate, rmin);
  block_reduce_rate_rho_once(rate, rmin, keys);
}

// The rate of the state `in` (nx, ny, nz interior cells) into out[0..1]: max rate,
// min density.  `out` doubles as the two ordered keys while the kernel runs.
template <bool FAST>
inline int launch_rate3d(const double* in, int64_t nx, int64_t ny, int64_t nz, double dx,
                         double dy, double dz, double b0x, double b0y, double b0z, double gamma,
                         double* out, hipStream_t s) {
  // 32-bit byte offsets within a plane of the potential
  if (nx < 2 || ny < 2 || nz < 2 || nx >= 65536 || (ny + 5) * (nz + 5) * 8 >= ((int64_t)1 << 32))
    return WB_EINVAL;
  StageArgs a{};
  a.in = in;
  a.g = Grid3{(int)nx + 4, (int)ny + 4, (int)nz + 4, (int)nx, (int)ny, (int)nz};
  a.dx = dx;
  a.dy = dy;
  a.dz = dz;
  a.b0x = b0x;
  a.b0y = b0y;
  a.b0z = b0z;
  a.gamma = gamma;
  const size_t X = nx + 4, Y = ny + 4, Z = nz + 4;
  a.off_ax = 5 * X * Y * Z;
  a.off_ay = a.off_ax + X * (Y + 1) * (Z + 1);
  a.off_az = a.off_ay + (X + 1) * Y * (Z + 1);
  unsigned long long* keys = (unsigned long long*)out;
  hipLaunchKernelGGL(k_rate_init, dim3(1), dim3(64), 0, s, keys);
  const dim3 grid((unsigned)((nz + 63) / 64), (unsigned)((ny + 3) / 4),
                  (unsigned)((nx + RATE_SEG - 1) / RATE_SEG));
  hipLaunchKernelGGL((k_rate3d_columns<FAST>), grid, dim3(256), 0, s, a, keys);
  hipLaunchKernelGGL(k_rate_finish, dim3(1), dim3(64), 0, s, keys, out);
  return wb_last_error();
}

}  // namespace wb
