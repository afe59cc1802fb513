// The fused 2D stage in FP64 (the parity path, -ffp-contract=off): the same bits
// as the split 2D kernels of wbmhd2d.hip.
#include "stage2d.hpp"
#include "stage2d_tiles.hpp"

using namespace wb;

extern "C" __attribute__((visibility("hidden"))) int wb_fill_stage_args2d(
    void* args, const double* in, double* out, int64_t nx, int64_t ny, double dx, double dy,
    double b0x, double b0y, double gamma, double eta, double kpar, double kperp, int limiter,
    const double* dt, int heun, int walls, int32_t* bad, int nxs) {
  if (nx < 2 || ny < 2 || nx > (1 << 20) || ny > (1 << 20) || limiter < 0 || limiter > 3 ||
      nxs < 1 || !(eta >= 0.0) || !(kpar >= 0.0) || !(kperp >= 0.0))
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
  // wall_mask (bit 0: x walls, bit 1: y walls) as the faces of the one patch
  a.walls = ((walls & 1) ? 3 : 0) | ((walls & 2) ? 12 : 0);
  a.diff = Diffusion{eta, kpar, kperp};
  return 0;
}

extern "C" int wb_mhd2d_stage(const double* in, double* out, int64_t nx, int64_t ny, double dx,
                              double dy, double b0x, double b0y, double gamma, double eta,
                              double kpar, double kperp, int limiter, const double* dt, int heun,
                              int wall_mask, int32_t* bad, int nxs, int tile, void* stream) {
  Stage2Args a;
  const int rc = wb_fill_stage_args2d(&a, in, out, nx, ny, dx, dy, b0x, b0y, gamma, eta, kpar,
                                      kperp, limiter, dt, heun, wall_mask, bad, nxs);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (a.diff.conduction() || eta != 0.0) {
    if (tile == 30) return launch_stage2d<double, 30, -1, true>(a, s);
    if (tile == 62) return launch_stage2d<double, 62, -1, true>(a, s);
    return launch_stage2d<double, 126, -1, true>(a, s);
  }
  if (tile == 30) return launch_stage2d<double, 30>(a, s);
  if (tile == 62) return launch_stage2d<double, 62>(a, s);
  if (tile == 254) return launch_stage2d<double, 254>(a, s);
  return launch_stage2d<double, 126>(a, s);
}

// All patches of a periodic 2D hierarchy in one persistent launch per stage, and the
// refluxing of the coarse cells next to finer patches.
template <int TY, bool NI>
static int launch_batched(const Stage2Args& a, const void* patches, const void* items, int nitems,
                          int32_t* counter, double* out, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL((k_stage2d_batched<double, TY, -1, NI>), dim3(grid), dim3(Row<TY>::THREADS),
                     0, s, a, (const PatchDesc*)patches, (const WorkItem*)items, nitems,
                     (int*)counter, out);
  return wb_last_error();
}

extern "C" int wb_amr2d_stage(const double* in, double* out, const void* patches,
                              const void* items, int nitems, int32_t* counter, double b0x,
                              double b0y, double gamma, double eta, double kpar, double kperp,
                              int limiter, const double* dt, int heun, int32_t* bad, int nxs,
                              int tile, int workgroups, void* stream) {
  Stage2Args a;
  const int rc = wb_fill_stage_args2d(&a, in, out, 2, 2, 1.0, 1.0, b0x, b0y, gamma, eta, kpar,
                                      kperp, limiter, dt, heun, 0, bad, nxs);
  if (rc) return rc;
  if (nitems <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, counter, 1);
  const unsigned grid = (unsigned)(workgroups > 0 ? workgroups : 1);
  const bool ni = a.diff.conduction() || eta != 0.0;
  if (tile == 30)
    return ni ? launch_batched<30, true>(a, patches, items, nitems, counter, out, grid, s)
              : launch_batched<30, false>(a, patches, items, nitems, counter, out, grid, s);
  if (tile == 62)
    return ni ? launch_batched<62, true>(a, patches, items, nitems, counter, out, grid, s)
              : launch_batched<62, false>(a, patches, items, nitems, counter, out, grid, s);
  return ni ? launch_batched<126, true>(a, patches, items, nitems, counter, out, grid, s)
            : launch_batched<126, false>(a, patches, items, nitems, counter, out, grid, s);
}

// The uniform grid's stage with a workgroup per 16 x 16 tile (ideal MHD; the halo is
// filled before the launch, as for wb_mhd2d_stage).
extern "C" int wb_mhd2d_stage_tiles(const double* in, double* out, int64_t nx, int64_t ny,
                                    double dx, double dy, double b0x, double b0y, double gamma,
                                    double eta, double kpar, double kperp, int limiter,
                                    const double* dt, int heun, int wall_mask, int32_t* bad,
                                    void* stream) {
  Stage2Args a;
  const int rc = wb_fill_stage_args2d(&a, in, out, nx, ny, dx, dy, b0x, b0y, gamma, eta, kpar,
                                      kperp, limiter, dt, heun, wall_mask, bad, 1);
  if (rc) return rc;
  const int tiles = (int)(((nx + 15) / 16) * ((ny + 15) / 16));
  return launch_stage2d_tiles<double, -1>(a, nullptr, nullptr, tiles, out, (hipStream_t)stream);
}

// Every 16 x 16 tile of every patch in one launch, a workgroup per tile (ideal MHD).
extern "C" int wb_amr2d_stage_tiles(const double* in, double* out, const void* patches,
                                    const void* items, int nitems, double b0x, double b0y,
                                    double gamma, double eta, double kpar, double kperp,
                                    int limiter, const double* dt, int heun, int32_t* bad,
                                    const int32_t* rec_first, const int32_t* rec_lists,
                                    double* rec, void* stream) {
  Stage2Args a;
  const int rc = wb_fill_stage_args2d(&a, in, out, 2, 2, 1.0, 1.0, b0x, b0y, gamma, eta, kpar,
                                      kperp, limiter, dt, heun, 0, bad, 1);
  if (rc) return rc;
  if (nitems <= 0) return 0;
  if (patches == nullptr || items == nullptr) return WB_EINVAL;
  if (rec_first != nullptr && (rec_lists == nullptr || rec == nullptr)) return WB_EINVAL;
  return launch_stage2d_tiles<double, -1>(a, patches, items, nitems, out, (hipStream_t)stream,
                                          rec_first, rec_lists, rec);
}

extern "C" int wb_amr2d_reflux_both(double* dst, const void* patches, const int32_t* rows,
                                    const int32_t* groups, int ngroups, const double* rec1,
                                    const double* rec2, const double* dt, void* stream) {
  return launch_reflux2d_both(dst, patches, rows, groups, ngroups, rec1, rec2, dt,
                              (hipStream_t)stream);
}

extern "C" int wb_amr2d_reflux(const double* src, double* dst, const void* patches,
                               const long long* rows, const int32_t* groups, int ngroups,
                               double b0x, double b0y, double gamma, double eta, double kpar,
                               double kperp, int limiter, const double* dt, int record,
                               double* reg, void* stream) {
  if (!(eta >= 0.0) || !(kpar >= 0.0) || !(kperp >= 0.0)) return WB_EINVAL;
  if (ngroups > 0 && reg == nullptr) return WB_EINVAL;
  if (ngroups <= 0) return 0;
  hipLaunchKernelGGL((k_reflux2d<double, -1>), dim3((unsigned)((ngroups + 63) / 64)), dim3(64), 0,
                     (hipStream_t)stream, src, dst, (const PatchDesc*)patches, rows,
                     (const int*)groups, ngroups, b0x, b0y, gamma, limiter,
                     Diffusion{eta, kpar, kperp}, dt, record, reg);
  return wb_last_error();
}
