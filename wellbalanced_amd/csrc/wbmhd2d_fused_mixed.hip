// The fused 2D stage with the reconstruction and the HLLD flux in FP32; the
// primitive conversion, the CT average, A_z and the update stay FP64.
#include "stage2d.hpp"

using namespace wb;

extern "C" __attribute__((visibility("hidden"))) int wb_fill_stage_args2d(void* args, const double* in, double* out, int64_t nx,
                                    int64_t ny, double dx, double dy, double b0x, double b0y,
                                    double gamma, int limiter, const double* dt, int heun,
                                    int walls, int32_t* bad, int nxs);

extern "C" int wb_mhd2d_stage_mixed(const double* in, double* out, int64_t nx, int64_t ny,
                                    double dx, double dy, double b0x, double b0y, double gamma,
                                    int limiter, const double* dt, int heun, int wall_mask,
                                    int32_t* bad, int nxs, int tile, void* stream) {
  Stage2Args a;
  const int rc = wb_fill_stage_args2d(&a, in, out, nx, ny, dx, dy, b0x, b0y, gamma, limiter, dt,
                                      heun, wall_mask, bad, nxs);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (tile == 62) return launch_stage2d<float, 62>(a, s);
  if (tile == 254) return launch_stage2d<float, 254>(a, s);
  return launch_stage2d<float, 126>(a, s);
}
