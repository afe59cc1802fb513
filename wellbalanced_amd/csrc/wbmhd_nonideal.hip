// The non-ideal 2D kernels as device kernels with the reference kernels' arguments
// (per-kernel twins): the resistive terms (kernels/mhd_resistive) and the
// anisotropic heat flux (kernels/mhd_conductive), added in place to the sweep's face
// fluxes and the CT-contact E_z.  FP64, -ffp-contract=off; the face formulas are the
// ones the fused stage kernels use.
//
// Layout as the twins of wbmhd2d.hip: prim (8, X, Y), bxf (X+1, Y), byf (X, Y+1),
// fx (7, X-3, Y) with x-face k+2 at fx[:, k], fy (7, X, Y-3) with y-face k+2 at
// fy[:, :, k], ez on the block's (X-3, Y-3) nodes.
#include "nonideal.hpp"
#include "wbmhd_common.hpp"

using namespace wb;

namespace {

__device__ __forceinline__ double jz_at(const double* bxf, const double* byf, int Y, int I, int J,
                                        double dx, double dy) {
  // bxf rows have Y entries, byf rows Y + 1 (mhd_resistive_pyccel.py:41-46)
  return node_current(byf[(size_t)I * (Y + 1) + J], byf[(size_t)(I - 1) * (Y + 1) + J],
                      bxf[(size_t)I * Y + J], bxf[(size_t)I * Y + J - 1], dx, dy);
}

// mhd_resistive_pyccel.py:47-67: part 0 the x-faces (f, j) of the block's rows, part
// 1 the y-faces (i, f) of its columns, part 2 the nodes.
__global__ void k_resistive2d(const double* __restrict__ prim, const double* __restrict__ bxf,
                              const double* __restrict__ byf, int X, int Y, double dx, double dy,
                              double eta, double* __restrict__ fx, double* __restrict__ fy,
                              double* __restrict__ ez) {
  const int h = 2, nx = X - 4, ny = Y - 4;
  const size_t PL = (size_t)X * Y;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int part = blockIdx.y;
  auto P = [&](int v, int i, int j) { return prim[v * PL + (size_t)i * Y + j]; };
  if (part == 0) {
    if (t >= (nx + 1) * ny) return;
    const int f = t / ny, i = h + f, j = h + t % ny;
    const size_t FX = (size_t)(X - 3) * Y, o = (size_t)f * Y + j;
    double e4 = fx[4 * FX + o], b6 = fx[6 * FX + o];
    resistive_xface(P(7, i - 1, j), P(7, i, j), P(6, i - 1, j), P(6, i, j),
                    jz_at(bxf, byf, Y, i, j, dx, dy), jz_at(bxf, byf, Y, i, j + 1, dx, dy), dx, eta,
                    e4, b6);
    fx[6 * FX + o] = b6;
    fx[4 * FX + o] = e4;
  } else if (part == 1) {
    if (t >= nx * (ny + 1)) return;
    const int i = h + t / (ny + 1), f = t % (ny + 1), j = h + f;
    const size_t FY = (size_t)X * (Y - 3), o = (size_t)i * (Y - 3) + f;
    double e4 = fy[4 * FY + o], b6 = fy[6 * FY + o];
    resistive_yface(P(7, i, j - 1), P(7, i, j), P(5, i, j - 1), P(5, i, j),
                    jz_at(bxf, byf, Y, i, j, dx, dy), jz_at(bxf, byf, Y, i + 1, j, dx, dy), dy, eta,
                    e4, b6);
    fy[6 * FY + o] = b6;
    fy[4 * FY + o] = e4;
  } else {
    if (t >= (nx + 1) * (ny + 1)) return;
    const int a = t / (ny + 1), b = t % (ny + 1);
    ez[t] = ez[t] + eta * jz_at(bxf, byf, Y, h + a, h + b, dx, dy);
  }
}

// mhd_conductive_pyccel.py:56-100: part 0 the x-faces (k, j), part 1 the y-faces
// (i, k) of the whole haloed block; the limited gradient along the face is zero in
// the block's first and last line.
__global__ void k_conductive2d(const double* __restrict__ prim, const double* __restrict__ bxf,
                               const double* __restrict__ byf, int X, int Y, double dx, double dy,
                               Diffusion d, double* __restrict__ fx, double* __restrict__ fy) {
  const size_t PL = (size_t)X * Y;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  auto P = [&](int v, int i, int j) { return prim[v * PL + (size_t)i * Y + j]; };
  auto T = [&](int i, int j) { return P(4, i, j) / P(0, i, j); };
  if (blockIdx.y == 0) {
    if (t >= (X - 3) * Y) return;
    const int k = t / Y, j = t % Y;
    auto along = [&](int i) {
      return (j >= 1 && j <= Y - 2) ? limited_gradient(T(i, j - 1), T(i, j), T(i, j + 1), dy) : 0.0;
    };
    const size_t FX = (size_t)(X - 3) * Y;
    double e4 = fx[4 * FX + t];
    conductive_face(bxf[(size_t)(k + 2) * Y + j], P(6, k + 1, j), P(6, k + 2, j), P(7, k + 1, j),
                    P(7, k + 2, j), T(k + 1, j), T(k + 2, j), along(k + 1), along(k + 2), dx, d, e4);
    fx[4 * FX + t] = e4;
  } else {
    if (t >= X * (Y - 3)) return;
    const int i = t / (Y - 3), k = t % (Y - 3);
    auto along = [&](int j) {
      return (i >= 1 && i <= X - 2) ? limited_gradient(T(i - 1, j), T(i, j), T(i + 1, j), dx) : 0.0;
    };
    const size_t FY = (size_t)X * (Y - 3);
    double e4 = fy[4 * FY + t];
    conductive_face(byf[(size_t)i * (Y + 1) + k + 2], P(5, i, k + 1), P(5, i, k + 2), P(7, i, k + 1),
                    P(7, i, k + 2), T(i, k + 1), T(i, k + 2), along(k + 1), along(k + 2), dy, d, e4);
    fy[4 * FY + t] = e4;
  }
}

}  // namespace

extern "C" {

int wb_resistive2d(const double* prim, const double* bxf, const double* byf, int64_t X, int64_t Y,
                   double dx, double dy, double eta, double* fx, double* fy, double* ez,
                   void* stream) {
  if (X < 5 || Y < 5 || X > (1 << 20) || Y > (1 << 20) || !(eta >= 0.0)) return WB_EINVAL;
  const int64_t most = (X - 3) * (Y - 3);
  dim3 grid((unsigned)((most + 255) / 256), 3);
  hipLaunchKernelGGL(k_resistive2d, grid, dim3(256), 0, (hipStream_t)stream, prim, bxf, byf, (int)X,
                     (int)Y, dx, dy, eta, fx, fy, ez);
  return wb_last_error();
}

int wb_conductive2d(const double* prim, const double* bxf, const double* byf, int64_t X, int64_t Y,
                    double dx, double dy, double parallel, double perpendicular, double* fx,
                    double* fy, void* stream) {
  if (X < 4 || Y < 4 || X > (1 << 20) || Y > (1 << 20) || !(parallel >= 0.0) ||
      !(perpendicular >= 0.0))
    return WB_EINVAL;
  const int64_t most = (X - 3) * Y > X * (Y - 3) ? (X - 3) * Y : X * (Y - 3);
  dim3 grid((unsigned)((most + 255) / 256), 2);
  hipLaunchKernelGGL(k_conductive2d, grid, dim3(256), 0, (hipStream_t)stream, prim, bxf, byf,
                     (int)X, (int)Y, dx, dy, Diffusion{0.0, parallel, perpendicular}, fx, fy);
  return wb_last_error();
}

}  // extern "C"
