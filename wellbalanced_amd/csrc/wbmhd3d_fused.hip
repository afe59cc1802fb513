// The fused 3D stage in FP64 (the parity path): built with -ffp-contract=off, it
// gives the same bits as the split kernels of wbmhd3d.hip.
#include "stage3d.hpp"

using namespace wb;

namespace {

size_t off3(int64_t nx, int64_t ny, int64_t nz, int which) {
  const size_t X = nx + 4, Y = ny + 4, Z = nz + 4;
  size_t off = 5 * X * Y * Z;
  if (which >= 1) off += X * (Y + 1) * (Z + 1);
  if (which >= 2) off += (X + 1) * Y * (Z + 1);
  return off;
}

}  // namespace

// Fills the kernel arguments shared by both precisions.
extern "C" __attribute__((visibility("hidden"))) int wb_fill_stage_args3d(void* args, const double* in, double* out, int64_t nx,
                                    int64_t ny, int64_t nz, double dx, double dy, double dz,
                                    double b0x, double b0y, double b0z, double gamma, int limiter,
                                    const double* dt, int heun, int32_t* bad, int nxs,
                                    int wrap);

int wb_fill_stage_args3d(void* args, const double* in, double* out, int64_t nx, int64_t ny,
                         int64_t nz, double dx, double dy, double dz, double b0x, double b0y,
                         double b0z, double gamma, int limiter, const double* dt, int heun,
                         int32_t* bad, int nxs, int wrap) {
  if (nx < 2 || ny < 2 || nz < 2 || nx > 4096 || ny > 4096 || nz > 4096 || limiter < 0 ||
      limiter > 3 || nxs < 1 || wrap < 0 || wrap > 7)
    return WB_EINVAL;
  StageArgs& a = *(StageArgs*)args;
  a.in = in;
  a.out = out;
  a.dt = dt;
  a.bad = (int*)bad;
  a.g = Grid3{(int)nx + 4, (int)ny + 4, (int)nz + 4, (int)nx, (int)ny, (int)nz};
  a.dx = dx;
  a.dy = dy;
  a.dz = dz;
  a.b0x = b0x;
  a.b0y = b0y;
  a.b0z = b0z;
  a.gamma = gamma;
  a.limiter = limiter;
  a.heun = heun;
  a.nxs = nxs;
  a.wrap = wrap;
  a.off_ax = off3(nx, ny, nz, 0);
  a.off_ay = off3(nx, ny, nz, 1);
  a.off_az = off3(nx, ny, nz, 2);
  return 0;
}

extern "C" int wb_mhd3d_stage(const double* in, double* out, int64_t nx, int64_t ny, int64_t nz,
                              double dx, double dy, double dz, double b0x, double b0y, double b0z,
                              double gamma, int limiter, const double* dt, int heun, int32_t* bad,
                              int nxs, int wrap, void* stream) {
  StageArgs a;
  const int rc = wb_fill_stage_args3d(&a, in, out, nx, ny, nz, dx, dy, dz, b0x, b0y, b0z, gamma,
                                      limiter, dt, heun, bad, nxs, wrap);
  if (rc) return rc;
  // two waves per SIMD (256 registers, 37 of them spilled; two workgroups' 78.9 KB of LDS
  // fit a CU): 4.4 ms per 256^3 launch against 6.7 ms at one wave without spills
  return launch_stage3d<double, 12, 16, 2>(a, (hipStream_t)stream);
}
