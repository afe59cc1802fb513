// The refined 3D hierarchy of equal blocks: device structures, the block-batched
// stage and the refluxing (templated on the flux arithmetic, instantiated by the
// FP64 and the mixed translation units).
//
// A block holds B^3 cells (B = 16) with a halo of 2 as the uniform 3D state stores
// a grid: 5 cell planes of 20^3, then A_x (20, 21, 21), A_y (21, 20, 21), A_z
// (21, 21, 20).  Level L + 1 refines level L by 2; lookup tables give the block at
// every block coordinate of a level.  The domain is periodic.
#pragma once

#include "rate3d.hpp"
#include "stage3d.hpp"

namespace wb {
namespace amr3 {

constexpr int B = 16;            // cells per block side
constexpr int X = B + 4;         // haloed cells per side
constexpr int XN = X + 1;        // haloed nodes per side
constexpr int NCELL = X * X * X;  // one cell field of a block
constexpr int NEDGE = X * XN * XN;  // one potential component of a block
constexpr int OFF_A = 5 * NCELL;    // first potential double of a block

struct BlockDesc {
  long long off;  // first double of the block in the buffer
  int level, bx, by, bz;
  double dx, dy, dz;
};

struct Levels {
  int nlev;
  int nb[6];    // blocks per axis
  int base[6];  // first entry of the level in the lookup table
};

inline bool levels_ok(const Levels& lv) { return lv.nlev >= 1 && lv.nlev <= 6; }

struct RefluxItem {
  int fine, coarse, ci, cj, ck, pad;
};

// Per coarse cell with coarse-fine faces: its n (1 to 3) faces as (item, lane) pairs.
struct RefluxCell {
  int n, item[3], lane[3], pad;
};

// Offsets inside a block: cell field f at (i, j, k); potential component c at
// (i, j, k) in its own extents (A_x (X, XN, XN), A_y (XN, X, XN), A_z (XN, XN, X)).
__device__ __forceinline__ int cell_at(int f, int i, int j, int k) {
  return f * NCELL + (i * X + j) * X + k;
}
__device__ __forceinline__ int edge_at(int c, int i, int j, int k) {
  if (c == 0) return OFF_A + (i * XN + j) * XN + k;
  if (c == 1) return OFF_A + NEDGE + (i * X + j) * XN + k;
  return OFF_A + 2 * NEDGE + (i * XN + j) * X + k;
}

// g modulo n in [0, n); coordinates within one period of the range (the fills' and
// the stencils' case) take no integer division.
__device__ __forceinline__ int wrap(int g, int n) {
  if (g >= 0 && g < n) return g;
  if (g < 0 && g >= -n) return g + n;
  if (g >= n && g < 2 * n) return g - n;
  g %= n;
  return g < 0 ? g + n : g;
}

// The block of `level` holding level cell/node coordinates (g0, g1, g2) canonically
// (the block whose interior range [0, B) contains each coordinate), or -1.
__device__ __forceinline__ int block_of(const int* lookup, const Levels& lv, int level, int g0,
                                        int g1, int g2) {
  const int nb = lv.nb[level];
  return lookup[lv.base[level] + ((g0 / B) * nb + g1 / B) * nb + g2 / B];
}

__device__ __forceinline__ StageArgs block_args(const StageArgs& base, const BlockDesc& d) {
  StageArgs a = base;
  a.in = base.in + d.off;
  a.out = base.out + d.off;
  a.dx = d.dx;
  a.dy = d.dy;
  a.dz = d.dz;
  return a;
}

// The offsets (doubles) of the 27 blocks of block d's level around it, index
// (x + 1) * 9 + (y + 1) * 3 + (z + 1), -1 where the level has no such block, into
// `hood` (LDS) for DIRECT reads; threads `t0` .. `t0` + 26 load it.
__device__ __forceinline__ void load_hood(long long* hood, const BlockDesc* __restrict__ blocks,
                                          const int* __restrict__ lookup, const Levels& lv,
                                          const BlockDesc& d, int t0 = 0) {
  const int t = threadIdx.x - t0;
  if (t >= 0 && t < 27) {
    const int nb = lv.nb[d.level];
    const int b0 = wrap(d.bx + t / 9 - 1, nb), b1 = wrap(d.by + (t / 3) % 3 - 1, nb),
              b2 = wrap(d.bz + t % 3 - 1, nb);
    const int id = lookup[lv.base[d.level] + (b0 * nb + b1) * nb + b2];
    hood[t] = id >= 0 ? blocks[id].off : -1;
  }
}

// Workgroups are dealt to the 8 XCDs in turn (workgroup w to XCD w % 8); slot `w / 8`
// of an XCD takes the next block of that XCD's contiguous eighth of the block list, so
// the blocks one XCD runs together are neighbours (the list is sorted by position
// within a level) and the cells they read around them stay in that XCD's L2.
__device__ __forceinline__ int xcd_block(int slot, int xcd, int nblocks) {
  const int chunk = (nblocks + 7) / 8;
  return xcd * chunk + slot;
}

// One stage of the listed blocks (ids; null: blocks 0, 1, ...): one workgroup per
// block, a 16 x 16 tile marching over the block's 16 planes.  DIRECT: the blocks read
// the cells around them in place through their neighbourhoods (load_hood).
template <typename P, int WAVES, int LIM, bool DIRECT>
__global__ void __launch_bounds__((Tile<B, B>::THREADS), WAVES)
k_stage3d_blocks(StageArgs base, const BlockDesc* __restrict__ blocks, const int* __restrict__ ids,
                 int nblocks, const int* __restrict__ lookup, Levels lv) {
  __shared__ StageShared<P, B, B> sh;
  __shared__ long long hood[27];
  const int k = xcd_block(blockIdx.x >> 3, blockIdx.x & 7, nblocks);
  if (k >= nblocks || (int)(blockIdx.x >> 3) >= (nblocks + 7) / 8) return;
  const int id = ids ? ids[k] : k;
  const BlockDesc d = blocks[id];
  if (DIRECT) {
    load_hood(hood, blocks, lookup, lv, d);
    __syncthreads();
  }
  stage3d_body<P, B, B, LIM, DIRECT>(block_args(base, d), 0, 0, 0, sh, hood, base.in,
                                     base.rec_slots ? base.rec_slots + 6 * id : nullptr);
}

// Half a block per workgroup: an 8 x 16 tile of 192 threads; slots 2 m and 2 m + 1 of
// an XCD are the two halves of its m-th block.  Three waves per SIMD (the
// register allocation spills a few values) run four such workgroups per CU, where one
// 384-thread workgroup per block fits only six waves at its natural 187 registers:
// 4.0 face solves per cell instead of 3.66, and the stage 30 % faster.  EARLY as in
// stage3d_body: the registers leave room for the early loads of the Heun average's
// edge values only, so the mixed path takes its cell values through LDS
// (LDS_OUT_CELLS | EARLY_OUT_EDGES; 7.7 KB more LDS, four workgroups still fit a CU).
template <typename P, int WAVES, int LIM, bool DIRECT, int EARLY = 0>
__global__ void __launch_bounds__((Tile<B / 2, B>::THREADS), WAVES)
k_stage3d_halves(StageArgs base, const BlockDesc* __restrict__ blocks, const int* __restrict__ ids,
                 int nblocks, const int* __restrict__ lookup, Levels lv) {
  __shared__ StageShared<P, B / 2, B> sh;
  __shared__ long long hood[27];
  __shared__ uint32_t dma[(EARLY & LDS_OUT_CELLS) ? 10 * Tile<B / 2, B>::THREADS : 1];
  const int slot = blockIdx.x >> 3, half = slot & 1;
  const int b = xcd_block(slot >> 1, blockIdx.x & 7, nblocks);
  if (b >= nblocks || (slot >> 1) >= (nblocks + 7) / 8) return;
  const int id = ids ? ids[b] : b;
  const BlockDesc d = blocks[id];
  if (DIRECT) {
    load_hood(hood, blocks, lookup, lv, d);
    __syncthreads();
  }
  stage3d_body<P, B / 2, B, LIM, DIRECT, EARLY>(block_args(base, d), half, 0, 0, sh, hood, base.in,
                                                base.rec_slots ? base.rec_slots + 6 * id : nullptr,
                                                dma);
}

// The CFL rate (mhd3d.py:674-724) over the cells of every block: a workgroup per
// block, each thread one (y, z) column marched along x as on the uniform grid (FAST:
// the mixed path's reciprocals), one atomic pair per workgroup.
// Two threads per column, each marching half the block (RATE_HALVES = 2 x 256 threads):
// twice the loads in flight of one thread per column.
constexpr int RATE_HALVES = 2;
template <bool FAST>
__global__ void __launch_bounds__(256 * RATE_HALVES) k_rate3d_blocks(
    StageArgs base, const BlockDesc* __restrict__ blocks, unsigned long long* __restrict__ keys) {
  const BlockDesc d = blocks[blockIdx.x];
  double rate = 0.0, rmin = __builtin_inf();
  const int col = threadIdx.x & 255, part = threadIdx.x >> 8;
  const int i0 = 2 + part * (B / RATE_HALVES);
  rate_column<FAST>(block_args(base, d), 2 + (col >> 4), 2 + (col & 15), i0,
                    i0 + B / RATE_HALVES, rate, rmin);
  block_reduce_rate_rho_once(rate, rmin, keys);
}

}  // namespace amr3
}  // namespace wb
