// The fused 3D stage with the reconstruction and the HLLD flux in FP32 and the
// primitive conversion, the cell electric field, the CT average, the potential
// and the update in FP64.
#include "rate3d.hpp"
#include "stage3d.hpp"

using namespace wb;

extern "C" __attribute__((visibility("hidden"))) int wb_fill_stage_args3d(void* args, const double* in, double* out, int64_t nx,
                                    int64_t ny, int64_t nz, double dx, double dy, double dz,
                                    double b0x, double b0y, double b0z, double gamma, int limiter,
                                    const double* dt, int heun, int32_t* bad, int nxs,
                                    int wrap);

extern "C" int wb_mhd3d_stage_mixed(const double* in, double* out, int64_t nx, int64_t ny,
                                    int64_t nz, double dx, double dy, double dz, double b0x,
                                    double b0y, double b0z, double gamma, int limiter,
                                    const double* dt, int heun, int32_t* bad, int nxs,
                                    int waves, int wrap, void* stream) {
  StageArgs a;
  const int rc = wb_fill_stage_args3d(&a, in, out, nx, ny, nz, dx, dy, dz, b0x, b0y, b0z, gamma,
                                      limiter, dt, heun, bad, nxs, wrap);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  return with_limiter(limiter, [&](auto lim) {
    constexpr int L = decltype(lim)::value;
    // at 3 waves per SIMD each stage takes its own early loads: stage 1 (no `out` read)
    // those of `in` only, stage 2 also the Heun average's cells through LDS and its edges
    // early (256^3 turbulence: 1292 and 1487 us against 1309 and 1499 with all of them
    // in registers)
    constexpr int IN = EARLY_IN_CELLS | EARLY_IN_EDGES;
    if (waves == 3 && heun)
      return launch_stage3d<float, 12, 16, 3, L, IN | LDS_OUT_CELLS | EARLY_OUT_EDGES>(a, s);
    if (waves == 3) return launch_stage3d<float, 12, 16, 3, L, IN>(a, s);
    return launch_stage3d<float, 12, 16, 2, L>(a, s);
  });
}

// The CFL rate of the mixed path: FP64, one division per cell, reciprocal spacings.
extern "C" int wb_rate3d_mixed(const double* state, int64_t nx, int64_t ny, int64_t nz, double dx,
                               double dy, double dz, double b0x, double b0y, double b0z,
                               double gamma, double* out, void* stream) {
  return launch_rate3d<true>(state, nx, ny, nz, dx, dy, dz, b0x, b0y, b0z, gamma, out,
                             (hipStream_t)stream);
}
