// The fused 2D stage with the reconstruction and the HLLD flux in FP32; the
// primitive conversion, the non-ideal face terms, the CT average, A_z and the update
// stay FP64.
#include "stage2d.hpp"
#include "stage2d_tiles.hpp"

using namespace wb;

extern "C" __attribute__((visibility("hidden"))) int wb_fill_stage_args2d(
    void* args, const double* in, double* out, int64_t nx, int64_t ny, double dx, double dy,
    double b0x, double b0y, double gamma, double eta, double kpar, double kperp, int limiter,
    const double* dt, int heun, int walls, int32_t* bad, int nxs);

extern "C" int wb_mhd2d_stage_mixed(const double* in, double* out, int64_t nx, int64_t ny,
                                    double dx, double dy, double b0x, double b0y, double gamma,
                                    double eta, double kpar, double kperp, int limiter,
                                    const double* dt, int heun, int wall_mask, int32_t* bad,
                                    int nxs, int tile, void* stream) {
  Stage2Args a;
  const int rc = wb_fill_stage_args2d(&a, in, out, nx, ny, dx, dy, b0x, b0y, gamma, eta, kpar,
                                      kperp, limiter, dt, heun, wall_mask, bad, nxs);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const bool ni = a.diff.conduction() || eta != 0.0;
  return with_limiter(limiter, [&](auto lim) {
    constexpr int L = decltype(lim)::value;
    if (ni) {
      if (tile == 30) return launch_stage2d<float, 30, L, true>(a, s);
      if (tile == 62) return launch_stage2d<float, 62, L, true>(a, s);
      return launch_stage2d<float, 126, L, true>(a, s);
    }
    if (tile == 30) return launch_stage2d<float, 30, L>(a, s);
    if (tile == 62) return launch_stage2d<float, 62, L>(a, s);
    if (tile == 254) return launch_stage2d<float, 254, L>(a, s);
    return launch_stage2d<float, 126, L>(a, s);
  });
}

// All patches of a periodic 2D hierarchy in one persistent launch per stage, and the
// refluxing of the coarse cells next to finer patches.
template <int TY, int L, bool NI>
static int launch_batched(const Stage2Args& a, const void* patches, const void* items, int nitems,
                          int32_t* counter, double* out, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL((k_stage2d_batched<float, TY, L, NI>), dim3(grid), dim3(Row<TY>::THREADS), 0,
                     s, a, (const PatchDesc*)patches, (const WorkItem*)items, nitems,
                     (int*)counter, out);
  return wb_last_error();
}

extern "C" int wb_amr2d_stage_mixed(const double* in, double* out, const void* patches,
                                    const void* items, int nitems, int32_t* counter, double b0x,
                                    double b0y, double gamma, double eta, double kpar,
                                    double kperp, int limiter, const double* dt, int heun,
                                    int32_t* bad, int nxs, int tile, int workgroups,
                                    void* stream) {
  Stage2Args a;
  const int rc = wb_fill_stage_args2d(&a, in, out, 2, 2, 1.0, 1.0, b0x, b0y, gamma, eta, kpar,
                                      kperp, limiter, dt, heun, 0, bad, nxs);
  if (rc) return rc;
  if (nitems <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, counter, 1);
  const unsigned grid = (unsigned)(workgroups > 0 ? workgroups : 1);
  const bool ni = a.diff.conduction() || eta != 0.0;
  return with_limiter(limiter, [&](auto lim) {
    constexpr int L = decltype(lim)::value;
    if (tile == 30)
      return ni ? launch_batched<30, L, true>(a, patches, items, nitems, counter, out, grid, s)
                : launch_batched<30, L, false>(a, patches, items, nitems, counter, out, grid, s);
    if (tile == 62)
      return ni ? launch_batched<62, L, true>(a, patches, items, nitems, counter, out, grid, s)
                : launch_batched<62, L, false>(a, patches, items, nitems, counter, out, grid, s);
    return ni ? launch_batched<126, L, true>(a, patches, items, nitems, counter, out, grid, s)
              : launch_batched<126, L, false>(a, patches, items, nitems, counter, out, grid, s);
  });
}

extern "C" int wb_mhd2d_stage_tiles_mixed(const double* in, double* out, int64_t nx, int64_t ny,
                                    double dx, double dy, double b0x, double b0y, double gamma,
                                    double eta, double kpar, double kperp, int limiter,
                                    const double* dt, int heun, int wall_mask, int32_t* bad,
                                    void* stream) {
  Stage2Args a;
  const int rc = wb_fill_stage_args2d(&a, in, out, nx, ny, dx, dy, b0x, b0y, gamma, eta, kpar,
                                      kperp, limiter, dt, heun, wall_mask, bad, 1);
  if (rc) return rc;
  const int tiles = (int)(((nx + 15) / 16) * ((ny + 15) / 16));
  hipStream_t s = (hipStream_t)stream;
  return with_limiter(limiter, [&](auto lim) {
    constexpr int L = decltype(lim)::value;
    return launch_stage2d_tiles<float, L>(a, nullptr, nullptr, tiles, out, s);
  });
}

extern "C" int wb_amr2d_stage_tiles_mixed(const double* in, double* out, const void* patches,
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
  hipStream_t s = (hipStream_t)stream;
  return with_limiter(limiter, [&](auto lim) {
    constexpr int L = decltype(lim)::value;
    return launch_stage2d_tiles<float, L>(a, patches, items, nitems, out, s, rec_first, rec_lists,
                                          rec);
  });
}

extern "C" int wb_amr2d_reflux_both_mixed(double* dst, const void* patches,
                                          const int32_t* rows, const int32_t* groups,
                                          int ngroups, const double* rec1, const double* rec2,
                                          const double* dt, void* stream) {
  return launch_reflux2d_both(dst, patches, rows, groups, ngroups, rec1, rec2, dt,
                              (hipStream_t)stream);
}

extern "C" int wb_amr2d_reflux_mixed(const double* src, double* dst, const void* patches,
                                     const long long* rows, const int32_t* groups, int ngroups,
                                     double b0x, double b0y, double gamma, double eta,
                                     double kpar, double kperp, int limiter, const double* dt,
                                     int record, double* reg, void* stream) {
  if (!(eta >= 0.0) || !(kpar >= 0.0) || !(kperp >= 0.0)) return WB_EINVAL;
  if (ngroups > 0 && reg == nullptr) return WB_EINVAL;
  if (ngroups <= 0) return 0;
  return with_limiter(limiter, [&](auto lim) {
    constexpr int L = decltype(lim)::value;
    hipLaunchKernelGGL((k_reflux2d<float, L>), dim3((unsigned)((ngroups + 63) / 64)), dim3(64), 0,
                       (hipStream_t)stream, src, dst, (const PatchDesc*)patches, rows,
                       (const int*)groups, ngroups, b0x, b0y, gamma, limiter,
                       Diffusion{eta, kpar, kperp}, dt, record, reg);
    return wb_last_error();
  });
}
