// The fused 2D stage in FP64 (the parity path, -ffp-contract=off): the same bits
// as the split 2D kernels of wbmhd2d.hip.
#include "stage2d.hpp"

using namespace wb;

extern "C" __attribute__((visibility("hidden"))) int wb_fill_stage_args2d(
    void* args, const double* in, double* out, int64_t nx, int64_t ny, double dx, double dy,
    double b0x, double b0y, double gamma, int limiter, const double* dt, int heun, int walls,
    int32_t* bad, int nxs) {
  if (nx < 2 || ny < 2 || nx > (1 << 20) || ny > (1 << 20) || limiter < 0 || limiter > 3 ||
      nxs < 1)
    return WB_EINVAL;
  Stage2Args& a = *(Stage2Args*)args;
  a.in = in;
  a.out = out;
  a.dt = dt;
  a.bad = (int*)bad;
  a.nx = (int)nx;
  a.ny = (int)ny;
  a.X = (int)nx + 4;
  a.Y = (int)ny + 4;
  a.dx = dx;
  a.dy = dy;
  a.b0x = b0x;
  a.b0y = b0y;
  a.gamma = gamma;
  a.limiter = limiter;
  a.heun = heun;
  a.nxs = nxs;
  a.walls = walls;
  return 0;
}

extern "C" int wb_mhd2d_stage(const double* in, double* out, int64_t nx, int64_t ny, double dx,
                              double dy, double b0x, double b0y, double gamma, int limiter,
                              const double* dt, int heun, int wall_mask, int32_t* bad, int nxs,
                              int tile, void* stream) {
  Stage2Args a;
  const int rc = wb_fill_stage_args2d(&a, in, out, nx, ny, dx, dy, b0x, b0y, gamma, limiter, dt,
                                      heun, wall_mask, bad, nxs);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (tile == 62) return launch_stage2d<double, 62>(a, s);
  if (tile == 254) return launch_stage2d<double, 254>(a, s);
  return launch_stage2d<double, 126>(a, s);
}
