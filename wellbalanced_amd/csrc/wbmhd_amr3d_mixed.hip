// The refined 3D hierarchy's stage with the FP32 reconstruction and HLLD (the mixed
// path, built with -ffp-contract=on); its refluxing reads the fluxes the stage recorded
// (wb_amr3d_reflux_both, shared with the FP64 path).
#include "amr3d.hpp"

using namespace wb;
using namespace wb::amr3;

extern "C" __attribute__((visibility("hidden"))) void wb_amr3d_base_args(
    void* args, const double* in, double* out, double b0x, double b0y, double b0z, double gamma,
    int limiter, const double* dt, int heun, int32_t* bad);

extern "C" {

int wb_amr3d_stage_mixed(const double* in, double* out, const void* blocks, const int32_t* ids,
                         int nids, int direct, const int32_t* lookup, const void* levels,
                         double b0x, double b0y, double b0z, double gamma, int limiter,
                         const double* dt, int heun, int32_t* bad, int waves,
                         const int32_t* rec_slots, double* rec, void* stream) {
  if (nids < 1 || (waves != 2 && waves != 3) || direct < 0 || direct > 1) return WB_EINVAL;
  if ((rec_slots == nullptr) != (rec == nullptr)) return WB_EINVAL;
  Levels lv{};
  if (direct) {
    if (!levels || !lookup) return WB_EINVAL;
    lv = *(const Levels*)levels;
    if (!levels_ok(lv)) return WB_EINVAL;
  }
  StageArgs a;
  wb_amr3d_base_args(&a, in, out, b0x, b0y, b0z, gamma, limiter, dt, heun, bad);
  a.rec_slots = (const int*)rec_slots;
  a.rec = rec;
  hipStream_t s = (hipStream_t)stream;
  const int* id = (const int*)ids;
  const int* lk = (const int*)lookup;
  return with_limiter(limiter, [&](auto lim) {
    constexpr int L = decltype(lim)::value;
    const dim3 halves(((nids + 7) / 8) * 16), tile_h(Tile<B / 2, B>::THREADS);
    const dim3 tile_b(Tile<B, B>::THREADS), blocks8(((nids + 7) / 8) * 8);
    // stage 2 takes the Heun average's cells through LDS and its edges early; stage 1,
    // which reads no `out`, runs fastest with the edges of `in` early (806 us, against
    // 814 with no early loads and 822 with stage 2's set)
    if (waves == 3 && direct && heun)
      hipLaunchKernelGGL((k_stage3d_halves<float, 3, L, true, LDS_OUT_CELLS | EARLY_OUT_EDGES>),
                         halves, tile_h, 0, s, a,
                         (const BlockDesc*)blocks, id, nids, lk, lv);
    else if (waves == 3 && direct)
      hipLaunchKernelGGL((k_stage3d_halves<float, 3, L, true, EARLY_IN_EDGES>), halves, tile_h,
                         0, s, a, (const BlockDesc*)blocks, id, nids, lk, lv);
    else if (waves == 3)
      hipLaunchKernelGGL((k_stage3d_halves<float, 3, L, false>), halves, tile_h, 0, s, a,
                         (const BlockDesc*)blocks, id, nids, lk, lv);
    else if (direct)
      hipLaunchKernelGGL((k_stage3d_blocks<float, 2, L, true>), blocks8, tile_b, 0, s, a,
                         (const BlockDesc*)blocks, id, nids, lk, lv);
    else
      hipLaunchKernelGGL((k_stage3d_blocks<float, 2, L, false>), blocks8, tile_b, 0, s, a,
                         (const BlockDesc*)blocks, id, nids, lk, lv);
    return wb_last_error();
  });
}

// The mixed path's CFL rate over every block: the column march of the uniform grid's
// mixed rate (one division per cell, reciprocals), in this FP32-contracted unit.
int wb_amr3d_rate_mixed(const double* buf, const void* blocks, int nblocks, double b0x,
                        double b0y, double b0z, double gamma, double* out, void* stream) {
  if (nblocks < 1) return WB_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  StageArgs a;
  wb_amr3d_base_args(&a, buf, const_cast<double*>(buf), b0x, b0y, b0z, gamma, 0, nullptr, 0,
                     nullptr);
  unsigned long long* keys = (unsigned long long*)out;
  hipLaunchKernelGGL(k_rate_init, dim3(1), dim3(64), 0, s, keys);
  hipLaunchKernelGGL(k_rate3d_blocks<true>, dim3((unsigned)nblocks), dim3(256 * RATE_HALVES), 0, s, a,
                     (const BlockDesc*)blocks, keys);
  hipLaunchKernelGGL(k_rate_finish, dim3(1), dim3(64), 0, s, keys, out);
  return wb_last_error();
}

}  // extern "C"
