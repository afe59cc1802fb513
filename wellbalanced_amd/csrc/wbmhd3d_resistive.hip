// The resistive terms of MHD3D (carafe mhd3d.py:561-602) as a correction after a
// stage: the stage kernels advance the ideal fluxes, then these kernels add what eta
// J and eta J x B change, from the same stage input.  FP64, -ffp-contract=off.
//
//   1. k_faces3d: the haloed face fields B = curl A + B0 (mhd3d.py:361-378);
//   2. k_current3d: J_c = d_a B_b - d_b B_a on the block's c-edges (mhd3d.py:578-590);
//      the potential of the stage output loses scale * dt * eta * J there;
//   3. k_energy3d: per block cell, the energy loses scale * sum_s dt/ds * (F_s+ - F_s-)
//      with F_s = eta (J_b B_c - J_c B_b) on its s-faces, J the mean of the face's two
//      edges along each direction and B the mean of the four faces around it
//      (mhd3d.py:591-602).
// scale is 1, or 0.5 for the Heun stage (the stage output is already the average).
// The ideal and resistive parts are summed in a different order than the reference's
// single flux difference: the results agree to round-off, not bit for bit.
#include <mutex>

#include "wbmhd_common.hpp"

namespace {

struct Grid {
  int X, Y, Z;          // haloed cells
  int nx, ny, nz;       // block cells
  size_t off_ax, off_ay, off_az;
  // workspace: face fields FX (X+1,Y,Z), FY (X,Y+1,Z), FZ (X,Y,Z+1), then J with the
  // extents of A: JX (X,Y+1,Z+1), JY (X+1,Y,Z+1), JZ (X+1,Y+1,Z)
  size_t fx, fy, fz, jx, jy, jz;
  __device__ size_t iFX(int i, int j, int k) const { return fx + ((size_t)i * Y + j) * Z + k; }
  __device__ size_t iFY(int i, int j, int k) const { return fy + ((size_t)i * (Y + 1) + j) * Z + k; }
  __device__ size_t iFZ(int i, int j, int k) const { return fz + ((size_t)i * Y + j) * (Z + 1) + k; }
  __device__ size_t iAX(int i, int j, int k) const { return ((size_t)i * (Y + 1) + j) * (Z + 1) + k; }
  __device__ size_t iAY(int i, int j, int k) const { return ((size_t)i * Y + j) * (Z + 1) + k; }
  __device__ size_t iAZ(int i, int j, int k) const { return ((size_t)i * (Y + 1) + j) * Z + k; }
};

Grid make_grid(int64_t nx, int64_t ny, int64_t nz) {
  Grid g;
  g.nx = (int)nx;
  g.ny = (int)ny;
  g.nz = (int)nz;
  g.X = g.nx + 4;
  g.Y = g.ny + 4;
  g.Z = g.nz + 4;
  const size_t X = g.X, Y = g.Y, Z = g.Z;
  g.off_ax = 5 * X * Y * Z;
  g.off_ay = g.off_ax + X * (Y + 1) * (Z + 1);
  g.off_az = g.off_ay + (X + 1) * Y * (Z + 1);
  g.fx = 0;
  g.fy = g.fx + (X + 1) * Y * Z;
  g.fz = g.fy + X * (Y + 1) * Z;
  g.jx = g.fz + X * Y * (Z + 1);
  g.jy = g.jx + X * (Y + 1) * (Z + 1);
  g.jz = g.jy + (X + 1) * Y * (Z + 1);
  return g;
}

size_t workspace_size(const Grid& g) {
  const size_t X = g.X, Y = g.Y, Z = g.Z;
  return g.jz + (X + 1) * (Y + 1) * Z;
}

// One thread per point of the (X+1)(Y+1)(Z+1) node lattice: the face fields that start
// there.
__global__ void k_faces3d(const double* __restrict__ in, Grid g, double dx, double dy, double dz,
                          double b0x, double b0y, double b0z, double* __restrict__ w) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)(g.X + 1) * (g.Y + 1) * (g.Z + 1);
  if (t >= n) return;
  const int k = (int)(t % (g.Z + 1)), j = (int)((t / (g.Z + 1)) % (g.Y + 1)),
            i = (int)(t / ((long long)(g.Z + 1) * (g.Y + 1)));
  const double* ax = in + g.off_ax;
  const double* ay = in + g.off_ay;
  const double* az = in + g.off_az;
  if (j < g.Y && k < g.Z)
    w[g.iFX(i, j, k)] = (az[g.iAZ(i, j + 1, k)] - az[g.iAZ(i, j, k)]) / dy -
                        (ay[g.iAY(i, j, k + 1)] - ay[g.iAY(i, j, k)]) / dz + b0x;
  if (i < g.X && k < g.Z)
    w[g.iFY(i, j, k)] = (ax[g.iAX(i, j, k + 1)] - ax[g.iAX(i, j, k)]) / dz -
                        (az[g.iAZ(i + 1, j, k)] - az[g.iAZ(i, j, k)]) / dx + b0y;
  if (i < g.X && j < g.Y)
    w[g.iFZ(i, j, k)] = (ay[g.iAY(i + 1, j, k)] - ay[g.iAY(i, j, k)]) / dx -
                        (ax[g.iAX(i, j + 1, k)] - ax[g.iAX(i, j, k)]) / dy + b0z;
}

// One thread per block node (I, J, K) in [2, 2+n]^3: the x-, y- and z-edges that start
// there and lie in the block; J into the workspace, the output potential updated.
__global__ void k_current3d(double* __restrict__ out, Grid g, double dx, double dy, double dz,
                            double eta, const double* __restrict__ dtp, int heun,
                            double* __restrict__ w) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mx = g.nx + 1, my = g.ny + 1, mz = g.nz + 1;
  if (t >= (long long)mx * my * mz) return;
  const int K = 2 + (int)(t % mz), J = 2 + (int)((t / mz) % my), I = 2 + (int)(t / ((long long)mz * my));
  const double f = (heun ? 0.5 : 1.0) * (*dtp) * eta;
  const double* F = w;  // the face fields, indexed by g.iFX / iFY / iFZ
  if (I < 2 + g.nx) {  // x-edge: cell I along x, nodes J, K
    const double jc = (F[g.iFZ(I, J, K)] - F[g.iFZ(I, J - 1, K)]) / dy -
                      (F[g.iFY(I, J, K)] - F[g.iFY(I, J, K - 1)]) / dz;
    w[g.jx + g.iAX(I, J, K)] = jc;
    double* a = out + g.off_ax + g.iAX(I, J, K);
    *a = *a - f * jc;
  }
  if (J < 2 + g.ny) {  // y-edge: nodes I, K, cell J
    const double jc = (F[g.iFX(I, J, K)] - F[g.iFX(I, J, K - 1)]) / dz -
                      (F[g.iFZ(I, J, K)] - F[g.iFZ(I - 1, J, K)]) / dx;
    w[g.jy + g.iAY(I, J, K)] = jc;
    double* a = out + g.off_ay + g.iAY(I, J, K);
    *a = *a - f * jc;
  }
  if (K < 2 + g.nz) {  // z-edge: nodes I, J, cell K
    const double jc = (F[g.iFY(I, J, K)] - F[g.iFY(I - 1, J, K)]) / dx -
                      (F[g.iFX(I, J, K)] - F[g.iFX(I, J - 1, K)]) / dy;
    w[g.jz + g.iAZ(I, J, K)] = jc;
    double* a = out + g.off_az + g.iAZ(I, J, K);
    *a = *a - f * jc;
  }
}

// The energy flux eta (J_b B_c - J_c B_b) through the s-face at node index I along s of
// the cell (i, j, k) (mhd3d.py:591-602).
__device__ double flux_x(const double* w, const Grid& g, int I, int j, int k, double eta) {
  const double jb = 0.5 * (w[g.jy + g.iAY(I, j, k)] + w[g.jy + g.iAY(I, j, k + 1)]);
  const double jc = 0.5 * (w[g.jz + g.iAZ(I, j, k)] + w[g.jz + g.iAZ(I, j + 1, k)]);
  const double bc = 0.5 * (0.5 * (w[g.iFZ(I - 1, j, k)] + w[g.iFZ(I, j, k)]) +
                           0.5 * (w[g.iFZ(I - 1, j, k + 1)] + w[g.iFZ(I, j, k + 1)]));
  const double bb = 0.5 * (0.5 * (w[g.iFY(I - 1, j, k)] + w[g.iFY(I, j, k)]) +
                           0.5 * (w[g.iFY(I - 1, j + 1, k)] + w[g.iFY(I, j + 1, k)]));
  return eta * (jb * bc - jc * bb);
}
__device__ double flux_y(const double* w, const Grid& g, int i, int J, int k, double eta) {
  const double jb = 0.5 * (w[g.jz + g.iAZ(i, J, k)] + w[g.jz + g.iAZ(i + 1, J, k)]);
  const double jc = 0.5 * (w[g.jx + g.iAX(i, J, k)] + w[g.jx + g.iAX(i, J, k + 1)]);
  const double bc = 0.5 * (0.5 * (w[g.iFX(i, J - 1, k)] + w[g.iFX(i, J, k)]) +
                           0.5 * (w[g.iFX(i + 1, J - 1, k)] + w[g.iFX(i + 1, J, k)]));
  const double bb = 0.5 * (0.5 * (w[g.iFZ(i, J - 1, k)] + w[g.iFZ(i, J, k)]) +
                           0.5 * (w[g.iFZ(i, J - 1, k + 1)] + w[g.iFZ(i, J, k + 1)]));
  return eta * (jb * bc - jc * bb);
}
__device__ double flux_z(const double* w, const Grid& g, int i, int j, int K, double eta) {
  const double jb = 0.5 * (w[g.jx + g.iAX(i, j, K)] + w[g.jx + g.iAX(i, j + 1, K)]);
  const double jc = 0.5 * (w[g.jy + g.iAY(i, j, K)] + w[g.jy + g.iAY(i + 1, j, K)]);
  const double bc = 0.5 * (0.5 * (w[g.iFY(i, j, K - 1)] + w[g.iFY(i, j, K)]) +
                           0.5 * (w[g.iFY(i, j + 1, K - 1)] + w[g.iFY(i, j + 1, K)]));
  const double bb = 0.5 * (0.5 * (w[g.iFX(i, j, K - 1)] + w[g.iFX(i, j, K)]) +
                           0.5 * (w[g.iFX(i + 1, j, K - 1)] + w[g.iFX(i + 1, j, K)]));
  return eta * (jb * bc - jc * bb);
}

__global__ void k_energy3d(double* __restrict__ out, Grid g, double dx, double dy, double dz,
                           double eta, const double* __restrict__ dtp, int heun,
                           const double* __restrict__ w) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx * g.ny * g.nz) return;
  const int k = 2 + (int)(t % g.nz), j = 2 + (int)((t / g.nz) % g.ny),
            i = 2 + (int)(t / ((long long)g.nz * g.ny));
  const double dt = *dtp;
  const double div = dt / dx * (flux_x(w, g, i + 1, j, k, eta) - flux_x(w, g, i, j, k, eta)) +
                     dt / dy * (flux_y(w, g, i, j + 1, k, eta) - flux_y(w, g, i, j, k, eta)) +
                     dt / dz * (flux_z(w, g, i, j, k + 1, eta) - flux_z(w, g, i, j, k, eta));
  double* e = out + 4 * (size_t)g.X * g.Y * g.Z + ((size_t)i * g.Y + j) * g.Z + k;
  *e = *e - (heun ? 0.5 * div : div);
}

inline unsigned grid_of(long long n) { return (unsigned)((n + 255) / 256); }


bool extents_ok(int64_t nx, int64_t ny, int64_t nz) {
  return nx >= 1 && ny >= 1 && nz >= 1 && nx <= 4096 && ny <= 4096 && nz <= 4096;
}

// The three passes on the workspace `work`; every workspace value they read is written
// by an earlier pass of the same call, so the workspace needs no clearing.
int launch_resistive(const double* in, double* out, int64_t nx, int64_t ny, int64_t nz,
                     double dx, double dy, double dz, double b0x, double b0y, double b0z,
                     double eta, const double* dt, int heun, double* work, hipStream_t s) {
  const Grid g = make_grid(nx, ny, nz);
  hipLaunchKernelGGL(k_faces3d, dim3(grid_of((long long)(g.X + 1) * (g.Y + 1) * (g.Z + 1))),
                     dim3(256), 0, s, in, g, dx, dy, dz, b0x, b0y, b0z, work);
  hipLaunchKernelGGL(k_current3d, dim3(grid_of((long long)(nx + 1) * (ny + 1) * (nz + 1))),
                     dim3(256), 0, s, out, g, dx, dy, dz, eta, dt, heun, work);
  hipLaunchKernelGGL(k_energy3d, dim3(grid_of(nx * ny * nz)), dim3(256), 0, s, out, g, dx, dy, dz,
                     eta, dt, heun, (const double*)work);
  return wb_last_error();
}

// The workspace of the allocation-free entry point: one array per GPU, held until the
// process ends, for grids whose workspace fits (up to 64^3 cells).  Host code never
// passes `scratch` itself (its host-side name is no device address): each device's
// address comes from hipGetSymbolAddress, taken once with that device current.
constexpr size_t SCRATCH_DOUBLES = size_t(1) << 21;  // 16 MiB
__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_lock;
double* scratch_address[MAX_DEVICES] = {};

int scratch_of_current_device(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return -(int)e;
  if (dev < 0 || dev >= MAX_DEVICES) return WB_EINVAL;
  std::lock_guard<std::mutex> hold(scratch_lock);
  if (scratch_address[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return -(int)e;
    scratch_address[dev] = (double*)p;
  }
  *out = scratch_address[dev];
  return 0;
}

}  // namespace

extern "C" {

int64_t wb_mhd3d_resistive_workspace(int64_t nx, int64_t ny, int64_t nz) {
  if (!extents_ok(nx, ny, nz)) return WB_EINVAL;
  return (int64_t)workspace_size(make_grid(nx, ny, nz));
}

int wb_mhd3d_resistive(const double* in, double* out, int64_t nx, int64_t ny, int64_t nz,
                       double dx, double dy, double dz, double b0x, double b0y, double b0z,
                       double eta, const double* dt, int heun, double* work, void* stream) {
  if (!extents_ok(nx, ny, nz) || !(eta >= 0.0)) return WB_EINVAL;
  if (eta == 0.0) return 0;
  return launch_resistive(in, out, nx, ny, nz, dx, dy, dz, b0x, b0y, b0z, eta, dt, heun, work,
                          (hipStream_t)stream);
}

int64_t wb_mhd3d_resistive_noalloc_limit(void) { return (int64_t)SCRATCH_DOUBLES; }

int wb_mhd3d_resistive_noalloc(const double* in, double* out, int64_t nx, int64_t ny,
                               int64_t nz, double dx, double dy, double dz, double b0x,
                               double b0y, double b0z, double eta, const double* dt, int heun,
                               void* stream) {
  if (!extents_ok(nx, ny, nz) || !(eta >= 0.0)) return WB_EINVAL;
  if (workspace_size(make_grid(nx, ny, nz)) > SCRATCH_DOUBLES) return WB_ESCRATCH;
  if (eta == 0.0) return 0;
  double* work = nullptr;
  const int rc = scratch_of_current_device(&work);
  if (rc) return rc;
  return launch_resistive(in, out, nx, ny, nz, dx, dy, dz, b0x, b0y, b0z, eta, dt, heun, work,
                          (hipStream_t)stream);
}

}  // extern "C"
