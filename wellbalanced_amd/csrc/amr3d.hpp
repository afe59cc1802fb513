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
  const BlockDesc d = blocks[ids ? ids[k] : k];
  if (DIRECT) {
    load_hood(hood, blocks, lookup, lv, d);
    __syncthreads();
  }
  stage3d_body<P, B, B, LIM, DIRECT>(block_args(base, d), 0, 0, 0, sh, hood, base.in);
}

// Half a block per workgroup: an 8 x 16 tile of 192 threads; slots 2 m and 2 m + 1 of
// an XCD are the two halves of its m-th block.  Three waves per SIMD (the
// register allocation spills a few values) run four such workgroups per CU, where one
// 384-thread workgroup per block fits only six waves at its natural 187 registers:
// 4.0 face solves per cell instead of 3.66, and the stage 30 % faster.
template <typename P, int WAVES, int LIM, bool DIRECT>
__global__ void __launch_bounds__((Tile<B / 2, B>::THREADS), WAVES)
k_stage3d_halves(StageArgs base, const BlockDesc* __restrict__ blocks, const int* __restrict__ ids,
                 int nblocks, const int* __restrict__ lookup, Levels lv) {
  __shared__ StageShared<P, B / 2, B> sh;
  __shared__ long long hood[27];
  const int slot = blockIdx.x >> 3, half = slot & 1;
  const int b = xcd_block(slot >> 1, blockIdx.x & 7, nblocks);
  if (b >= nblocks || (slot >> 1) >= (nblocks + 7) / 8) return;
  const BlockDesc d = blocks[ids ? ids[b] : b];
  if (DIRECT) {
    load_hood(hood, blocks, lookup, lv, d);
    __syncthreads();
  }
  stage3d_body<P, B / 2, B, LIM, DIRECT>(block_args(base, d), half, 0, 0, sh, hood, base.in);
}

// The flux through the low `axis` face of haloed cell (i, j, k) of the block state
// `a.in`, computed as the stage kernel computes it: the four cells along the normal
// by CellPrim, the frame rows of the sweep, the face field of the potential, then
// the face solve of the path (rows of the normal frame, as double).
// `hood` (with `root`): the block's 27-block neighbourhood (load_hood), through which a
// cell outside the block's interior is read in place from the block of its level
// holding it, or, where the level has none, from the block's own (filled) ghosts; null:
// every cell from the block's ghosts.
template <typename P, int LIM>
__device__ void face_flux3d(const StageArgs& a, int axis, int i, int j, int k, double out[7],
                            const long long* hood = nullptr, const double* root = nullptr) {
  constexpr bool FAST = sizeof(P) == 4;
  const Grid3 g = a.g;
  const EdgePot A{a.in + a.off_ax, a.in + a.off_ay, a.in + a.off_az, g.Y, g.Z};
  const Addr3 d = addr3(a);
  const double rdx = 1.0 / a.dx, rdy = 1.0 / a.dy, rdz = 1.0 / a.dz;
  constexpr int cx[7] = {0, 1, 2, 3, 4, 6, 7};
  constexpr int cy[7] = {0, 2, 3, 1, 4, 7, 5};
  constexpr int cz[7] = {0, 3, 1, 2, 4, 5, 6};
  P q[7][4];
  for (int s = 0; s < 4; ++s) {
    const int ii = i + (axis == 0 ? s - 2 : 0), jj = j + (axis == 1 ? s - 2 : 0),
              kk = k + (axis == 2 ? s - 2 : 0);
    const int ri = ii < 2 ? -1 : (ii >= B + 2 ? 1 : 0), rj = jj < 2 ? -1 : (jj >= B + 2 ? 1 : 0),
              rk = kk < 2 ? -1 : (kk >= B + 2 ? 1 : 0);
    const long long off = hood ? hood[(ri + 1) * 9 + (rj + 1) * 3 + rk + 1] : -1;
    const bool there = off >= 0 && (ri | rj | rk) != 0;
    const int pi = there ? ii - ri * B : ii, pj = there ? jj - rj * B : jj,
              pk = there ? kk - rk * B : kk;
    const char* src = there ? reinterpret_cast<const char*>(root + off)
                            : reinterpret_cast<const char*>(a.in);
    CellPrim<FAST> c;
    c.load(a, d, src, pi, (uint32_t)pj * d.crow + (uint32_t)pk * 8u,
           (uint32_t)pj * d.arow + (uint32_t)pk * 8u, rdx, rdy, rdz);
    const P pr[8] = {P(c.r), P(c.u), P(c.v), P(c.w), P(c.p), P(c.bx), P(c.by), P(c.bz)};
    for (int r = 0; r < 7; ++r) q[r][s] = pr[axis == 0 ? cx[r] : (axis == 1 ? cy[r] : cz[r])];
  }
  double bn;
  if (axis == 0) {
    bn = FAST ? (A.AZ(i, j + 1, k) - A.AZ(i, j, k)) * rdy - (A.AY(i, j, k + 1) - A.AY(i, j, k)) * rdz + a.b0x
              : A.bx(i, j, k, a.dy, a.dz, a.b0x);
  } else if (axis == 1) {
    bn = FAST ? (A.AX(i, j, k + 1) - A.AX(i, j, k)) * rdz - (A.AZ(i + 1, j, k) - A.AZ(i, j, k)) * rdx + a.b0y
              : A.by(i, j, k, a.dz, a.dx, a.b0y);
  } else {
    bn = FAST ? (A.AY(i + 1, j, k) - A.AY(i, j, k)) * rdx - (A.AX(i, j + 1, k) - A.AX(i, j, k)) * rdy + a.b0z
              : A.bz(i, j, k, a.dx, a.dy, a.b0z);
  }
  P fl[7];
  face_solve<P, LIM>(q, a.limiter, P(bn), P(a.gamma), P(1.0 / (a.gamma - 1.0)), fl);
  for (int r = 0; r < 7; ++r) out[r] = double(fl[r]);
}

// The flux register and refluxing (FluxRegister.update with scale 0.5 dt and
// FluxRegister.reflux as MHD3D.step calls them, mhd3d.py:650-655,745-750): per
// coarse-fine block face, the 8 x 8 coarse cells across it trade the flux of their face
// on the fine side for the mean of the four fine fluxes, both recomputed from the stage
// input `src`: d = sum over the cell's faces of dt/h (F_c - mean) or dt/h (mean - F_c).
// record (after stage 1): reg[5 t + f] = 0.5 d; otherwise (after stage 2 and its Heun
// average) the cell gains reg + 0.5 d.  An item is one such face (dir = 2 axis + side;
// side 0: the coarse cells at the fine block's low face, their +axis face; side 1: at
// its high face).
//
// A workgroup of 256 threads per REFLUX_CELLS = 51 coarse cells that have such faces:
// thread 5 m + k takes flux k of cell m, k = 0 the coarse flux of the cell's face and
// k = 1..4 the four fine fluxes (p, q) = ((k - 1) / 2, (k - 1) % 2); the fine ones meet
// the coarse one's thread through LDS (four waves, so that two workgroups fill the 8
// waves a CU holds at this kernel's register count).  `cells` lists, per cell, its n
// (1 to 3) faces as (item, lane) pairs in increasing direction order, the lane u * 8 + v
// naming the cell inside the item's 8 x 8 face (the host puts the cells with more faces
// first, so that a workgroup's cells share one face count).  The coarse flux's thread
// sums the corrections of a cell's faces in that order, with no atomics on the state
// (deterministic).  `hoods`: per block
// the 27 offsets of the blocks of its level around it (-1: none) for reads in place, or
// null to read the filled ghosts.
constexpr int REFLUX_CELLS = 51;

struct RefluxCell {
  int n, item[3], lane[3], pad;
};

template <typename P, int LIM>
__global__ void __launch_bounds__(256) k_reflux3d(StageArgs base, const BlockDesc* __restrict__ blocks,
                                                  const RefluxItem* __restrict__ items,
                                                  const RefluxCell* __restrict__ cells, int ncells,
                                                  const long long* __restrict__ hoods,
                                                  double* __restrict__ dst, int record,
                                                  double* __restrict__ reg) {
  __shared__ double fine[4][7][REFLUX_CELLS];
  __shared__ int most;
  const int m = threadIdx.x / 5, kf = threadIdx.x % 5;
  const int t = blockIdx.x * REFLUX_CELLS + m;
  const int n = m < REFLUX_CELLS && t < ncells ? cells[t].n : 0;
  if (threadIdx.x == 0) most = 0;
  __syncthreads();
  if (kf == 0 && n > 0) atomicMax(&most, n);
  __syncthreads();
  const int nmax = most;
  const double dt = *base.dt;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int c[3] = {0, 0, 0};
  BlockDesc cd{};
  for (int f = 0; f < nmax; ++f) {
    const bool has = f < n;
    int axis = 0, side = 0;
    double fl[7];
    if (has) {
      const RefluxItem item = items[cells[t].item[f]];
      const int lc = cells[t].lane[f];
      axis = item.pad >> 1;
      side = item.pad & 1;
      const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
      const int u = lc >> 3, v = lc & 7;
      // the coarse cell (haloed local) and the face of it on the fine side
      c[0] = item.ci + 2;
      c[1] = item.cj + 2;
      c[2] = item.ck + 2;
      c[a1] += u;
      c[a2] += v;
      int g[3];
      if (kf == 0) {
        g[0] = c[0];
        g[1] = c[1];
        g[2] = c[2];
        if (side == 0) g[axis] += 1;  // its +axis face is the low face of the next cell
      } else {  // fine face (p, q): the fine block's low face (side 0) or high face
        g[axis] = side == 0 ? 2 : 2 + B;
        g[a1] = 2 + 2 * u + ((kf - 1) >> 1);
        g[a2] = 2 + 2 * v + ((kf - 1) & 1);
      }
      const int blk = kf == 0 ? item.coarse : item.fine;
      const BlockDesc bd = blocks[blk];
      cd = bd;
      face_flux3d<P, LIM>(block_args(base, bd), axis, g[0], g[1], g[2], fl,
                          hoods ? hoods + 27 * (size_t)blk : nullptr, base.in);
      if (kf > 0)
        for (int r = 0; r < 7; ++r) fine[kf - 1][r][m] = fl[r];
    }
    __syncthreads();
    if (has && kf == 0) {
      const double dtd = axis == 0 ? dt / cd.dx : (axis == 1 ? dt / cd.dy : dt / cd.dz);
      constexpr int xrow[5] = {0, 1, 2, 3, 4};
      constexpr int yrow[5] = {0, 3, 1, 2, 4};
      constexpr int zrow[5] = {0, 2, 3, 1, 4};
      for (int fld = 0; fld < 5; ++fld) {
        const int row = axis == 0 ? xrow[fld] : (axis == 1 ? yrow[fld] : zrow[fld]);
        const double fm =
            ((fine[0][row][m] + fine[1][row][m]) + (fine[2][row][m] + fine[3][row][m])) * 0.25;
        const double d = side == 0 ? dtd * (fl[row] - fm) : dtd * (fm - fl[row]);
        acc[fld] = acc[fld] + d;
      }
    }
    __syncthreads();
  }
  if (n == 0 || kf != 0) return;
  double* kept = reg + 5 * (size_t)t;
  if (record) {
    for (int fld = 0; fld < 5; ++fld) kept[fld] = 0.5 * acc[fld];
    return;
  }
  double* out = dst + cd.off;
  for (int fld = 0; fld < 5; ++fld) {
    const int o = cell_at(fld, c[0], c[1], c[2]);
    out[o] = out[o] + (kept[fld] + 0.5 * acc[fld]);
  }
}

// The CFL rate (mhd3d.py:674-724) over the cells of every block: a workgroup per
// block, each thread one (y, z) column marched along x as on the uniform grid (FAST:
// the mixed path's reciprocals), one atomic pair per workgroup.
template <bool FAST>
__global__ void __launch_bounds__(256) k_rate3d_blocks(StageArgs base,
                                                       const BlockDesc* __restrict__ blocks,
                                                       unsigned long long* __restrict__ keys) {
  const BlockDesc d = blocks[blockIdx.x];
  double rate = 0.0, rmin = __builtin_inf();
  rate_column<FAST>(block_args(base, d), 2 + (threadIdx.x >> 4), 2 + (threadIdx.x & 15), 2,
                    2 + B, rate, rmin);
  block_reduce_rate_rho_once(rate, rmin, keys);
}

}  // namespace amr3
}  // namespace wb
