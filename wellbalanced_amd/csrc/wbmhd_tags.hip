// The reference's refinement estimator, current_tags: cells where the current
// density times the cell size exceeds a threshold.  The expressions are the
// reference's (numpy order), so the tags equal its tags bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wbmhd_common.hpp"

namespace {

// 2D (mhd.py:1011-1041 with _current, mhd.py:826-831): J_z at node (I, J) of the
// haloed block from the face fields bxf (X+1, Y) and byf (X, Y+1) of the primitive
// kernel; a cell's |J_z| is the mean of |J_z| at its four nodes.
__device__ __forceinline__ double jz_node(const double* bxf, const double* byf, int Y, int I, int J,
                                          double dx, double dy) {
  return (byf[(size_t)I * (Y + 1) + J] - byf[(size_t)(I - 1) * (Y + 1) + J]) / dx -
         (bxf[(size_t)I * Y + J] - bxf[(size_t)I * Y + J - 1]) / dy;
}

__global__ void k_current_tags2d(const double* __restrict__ bxf, const double* __restrict__ byf,
                                 int Y, int nx, int ny, double dx, double dy, double threshold,
                                 uint8_t* __restrict__ out) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nx * ny) return;
  const int i = id / ny, j = id % ny;
  const int I = i + 2, J = j + 2;
  double cell = 0.0;
  cell = cell + __builtin_fabs(jz_node(bxf, byf, Y, I, J, dx, dy));
  cell = cell + __builtin_fabs(jz_node(bxf, byf, Y, I, J + 1, dx, dy));
  cell = cell + __builtin_fabs(jz_node(bxf, byf, Y, I + 1, J, dx, dy));
  cell = cell + __builtin_fabs(jz_node(bxf, byf, Y, I + 1, J + 1, dx, dy));
  cell = 0.25 * cell;
  const double h = dx > dy ? dx : dy;
  out[id] = cell * h > threshold ? 1 : 0;
}

// 3D (mhd3d.py:771-801): J_c = dB_b/dx_a - dB_a/dx_b, (c, a, b) cyclic, by central
// differences (f[+1] - f[-1]) / (2 h) of the cell field prim[5..7] (8, X, Y, Z).
__global__ void k_current_tags3d(const double* __restrict__ prim, int X, int Y, int Z, int nx,
                                 int ny, int nz, double dx, double dy, double dz, double threshold,
                                 uint8_t* __restrict__ out) {
  const size_t n = (size_t)nx * ny * nz;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const int k = id % nz + 2, t = id / nz, j = t % ny + 2, i = t / ny + 2;
  const size_t P = (size_t)X * Y * Z;
  const size_t sx = (size_t)Y * Z, sy = Z, c0 = ((size_t)i * Y + j) * Z + k;
  const double* B[3] = {prim + 5 * P, prim + 6 * P, prim + 7 * P};
  const size_t stride[3] = {sx, sy, 1};
  const double h[3] = {dx, dy, dz};
  double square = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int a = (c + 1) % 3, b = (c + 2) % 3;
    const double ga = (B[b][c0 + stride[a]] - B[b][c0 - stride[a]]) / (2.0 * h[a]);
    const double gb = (B[a][c0 + stride[b]] - B[a][c0 - stride[b]]) / (2.0 * h[b]);
    const double jc = ga - gb;
    square = square + jc * jc;
  }
  double hm = dx > dy ? dx : dy;
  hm = hm > dz ? hm : dz;
  out[id] = __builtin_sqrt(square) * hm > threshold ? 1 : 0;
}

}  // namespace

extern "C" {

int wb_current_tags2d(const double* bxf, const double* byf, int64_t nx, int64_t ny, double dx,
                      double dy, double threshold, uint8_t* out, void* stream) {
  if (nx < 1 || ny < 1 || nx * ny > (int64_t)1 << 30) return WB_EINVAL;
  const int n = (int)(nx * ny);
  hipLaunchKernelGGL(k_current_tags2d, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     bxf, byf, (int)ny + 4, (int)nx, (int)ny, dx, dy, threshold, out);
  return wb_last_error();
}

int wb_current_tags3d(const double* prim, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                      double dz, double threshold, uint8_t* out, void* stream) {
  if (nx < 1 || ny < 1 || nz < 1 || nx * ny * nz > (int64_t)1 << 31) return WB_EINVAL;
  const size_t n = (size_t)(nx * ny * nz);
  hipLaunchKernelGGL(k_current_tags3d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, prim, (int)nx + 4, (int)ny + 4, (int)nz + 4, (int)nx,
                     (int)ny, (int)nz, dx, dy, dz, threshold, out);
  return wb_last_error();
}

}  // extern "C"
