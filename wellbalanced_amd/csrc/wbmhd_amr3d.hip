// The refined 3D hierarchy: ghost fill (whole shells, or the regions a block's level
// has no block for), average-down and restriction of the potential, the FP64 CFL rate
// over every block, and the FP64 stage and refluxing (the parity path, built with
// -ffp-contract=off).
#include "amr3d.hpp"

using namespace wb;
using namespace wb::amr3;

namespace {

// Workgroups per block that fill its cells (each a contiguous quarter of the shell);
// three more fill the three potential components.
constexpr int CELL_PARTS = 4;
// Workgroups per refined block and potential component of the restriction.
constexpr int RESTRICT_PARTS = 4;

// The t-th position of the ghost shell of an array of extents e (the positions with a
// coordinate outside the block's own range [2, 2 + B)): the slab of the low and high
// x layers, then the y layers inside it, then the z layers; shell_size counts them.
__device__ __forceinline__ int shell_size(int e0, int e1, int e2) {
  return (e0 - B) * e1 * e2 + B * (e1 - B) * e2 + B * B * (e2 - B);
}
__device__ __forceinline__ void shell_at(int t, int e0, int e1, int e2, int p[3]) {
  const int n0 = (e0 - B) * e1 * e2, n1 = B * (e1 - B) * e2;
  if (t < n0) {
    const int o = t / (e1 * e2);
    p[0] = o < 2 ? o : o + B;
    p[1] = (t / e2) % e1;
    p[2] = t % e2;
  } else if (t < n0 + n1) {
    t -= n0;
    const int w = (e1 - B) * e2;
    p[0] = 2 + t / w;
    const int r = t % w, o = r / e2;
    p[1] = o < 2 ? o : o + B;
    p[2] = r % e2;
  } else {
    t -= n0 + n1;
    const int w = B * (e2 - B);
    p[0] = 2 + t / w;
    const int r = t % w, o = r % (e2 - B);
    p[1] = 2 + r / (e2 - B);
    p[2] = o < 2 ? o : o + B;
  }
}

// The 27 blocks around a block on its own level, or around its parent on the level
// below (relative block offsets -1..1 per axis, x slowest): ids (-1: none) and first
// doubles.  Every source of a block's ghost fill lies in one of the two, so a fill
// workgroup loads them once into LDS and needs no level lookup per element.
struct Hood {
  long long off[27];
  int id[27];
};

__device__ __forceinline__ int hood_index(int r0, int r1, int r2) {
  return ((r0 + 1) * 3 + (r1 + 1)) * 3 + (r2 + 1);
}

// floor(u / B) for u in [-B, 2B).
__device__ __forceinline__ int rel_of(int u) { return u < 0 ? -1 : (u >= B ? 1 : 0); }

// The master copy of an edge (component c) at offsets u from the origin of the block at
// the centre of h, in that level's cells: the block holding it in its own range [0, B)
// first, then the blocks holding it on their upper boundary (local index B on a node
// axis where the coordinate is a multiple of B), in a fixed order (bit s of the
// candidate: x, y, z from the top); returns its index in h (-1: none) and its local
// haloed coordinates.  The fills never write a master copy.
__device__ __forceinline__ int hood_master(const Hood& h, int c, const int u[3], int loc[3]) {
  int rel[3], um[3];
  for (int a = 0; a < 3; ++a) {
    rel[a] = rel_of(u[a]);
    um[a] = u[a] - rel[a] * B;
  }
  for (int s = 0; s < 8; ++s) {
    bool ok = true;
    int r[3];
    for (int a = 0; a < 3; ++a) {
      const int sa = (s >> (2 - a)) & 1;
      if (sa && (a == c || um[a] != 0)) ok = false;
      r[a] = rel[a] - sa;
      if (r[a] < -1) ok = false;
      loc[a] = sa ? B + 2 : um[a] + 2;
    }
    if (!ok) continue;
    const int k = hood_index(r[0], r[1], r[2]);
    if (h.id[k] >= 0) return k;
  }
  return -1;
}

// Edge p (haloed, component C) of the block at `me` outside its own range: unless
// `all`, the value of its master copy on the same level (none written if the block
// holds the master itself); without one (or with `all`), interpolated from the level
// below: along the cell axis by the weights 1/4, 3/4, across it the coarse node (even
// fine nodes) or the mean of the two coarse nodes (odd).
template <int C>
__device__ __forceinline__ void fill_edge_at(double* __restrict__ buf, const Hood& same,
                                             const Hood& coarse, bool all, const int bo[3],
                                             const int po[3], double* __restrict__ me,
                                             const int p[3]) {
  constexpr int c = C;
  if (!all) {
    const int u[3] = {p[0] - 2, p[1] - 2, p[2] - 2};
    int loc[3];
    const int k = hood_master(same, c, u, loc);
    if (k >= 0) {
      if (k == 13 && loc[0] == p[0] && loc[1] == p[1] && loc[2] == p[2])
        return;  // this block's copy is the master
      me[edge_at(c, p[0], p[1], p[2])] = buf[same.off[k] + edge_at(c, loc[0], loc[1], loc[2])];
      return;
    }
  }
  int c0[3], nt[3];
  double w[3][2];
  for (int a = 0; a < 3; ++a) {
    const int g = bo[a] + p[a] - 2;  // level coordinate (unwrapped: the hood wraps)
    const int m = g >> 1;
    if (a == c) {  // cell axis
      const double tt = (g & 1) ? 0.25 : 0.75;
      c0[a] = (g & 1) ? m : m - 1;
      nt[a] = 2;
      w[a][0] = 1.0 - tt;
      w[a][1] = tt;
    } else if ((g & 1) == 0) {  // a coarse node
      c0[a] = m;
      nt[a] = 1;
      w[a][0] = 1.0;
      w[a][1] = 0.0;
    } else {  // the midpoint of coarse nodes m and m + 1
      c0[a] = m;
      nt[a] = 2;
      w[a][0] = 0.5;
      w[a][1] = 0.5;
    }
  }
  // the (up to) 8 sources first, then their loads together, then the sum in source
  // order over the ones that exist (a source outside the level below is excluded by
  // the nesting rule; its load reads a valid dummy address and is not summed)
  long long src[8];
  bool ok[8];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int z = 0; z < 2; ++z) {
        const int i = (x * 2 + y) * 2 + z;
        const int u[3] = {c0[0] + x - po[0], c0[1] + y - po[1], c0[2] + z - po[2]};
        int cl[3];
        const bool used = x < nt[0] && y < nt[1] && z < nt[2];
        const int k = used ? hood_master(coarse, c, u, cl) : -1;
        ok[i] = k >= 0;
        src[i] = ok[i] ? coarse.off[k] + edge_at(c, cl[0], cl[1], cl[2]) : 0;
      }
  double v[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = buf[src[i]];
  double acc = 0.0;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int z = 0; z < 2; ++z) {
        const int i = (x * 2 + y) * 2 + z;
        const double t = acc + ((w[0][x] * w[1][y]) * w[2][z]) * v[i];
        acc = ok[i] ? t : acc;
      }
  me[edge_at(c, p[0], p[1], p[2])] = acc;
}

// The ghost edges of component C of one block (see k_fill3d): extents and shell
// divisions are compile-time constants per component.
template <int C>
__device__ __forceinline__ void fill_edges(double* __restrict__ buf, const Hood& same,
                                           const Hood& coarse, bool all, bool restricted,
                                           const int bo[3], const int po[3],
                                           double* __restrict__ me) {
  constexpr int c = C;
  constexpr int e0 = c == 0 ? X : XN, e1 = c == 1 ? X : XN, e2 = c == 2 ? X : XN;
  const int nedge = all ? e0 * e1 * e2 : shell_size(e0, e1, e2);
  for (int t = threadIdx.x; t < nedge; t += blockDim.x) {
    int p[3];
    if (all) {
      p[0] = t / (e1 * e2);
      p[1] = (t / e2) % e1;
      p[2] = t % e2;
    } else {
      shell_at(t, e0, e1, e2, p);
    }
    if (restricted) {
      // only an edge of a ghost cell whose level has no block holding it (the cells
      // with one are read in place by the stage): c is the edge's cell axis, the
      // others node axes shared by the cells at p - 1 and p
      bool needed = false;
      for (int s = 0; s < 4 && !needed; ++s) {
        int q[3];
        bool ok = true;
        for (int a = 0, bit = 0; a < 3; ++a) {
          q[a] = p[a];
          if (a != c) q[a] -= (s >> bit++) & 1;
          ok = ok && q[a] >= 0 && q[a] < X;
        }
        if (!ok || (q[0] >= 2 && q[0] < B + 2 && q[1] >= 2 && q[1] < B + 2 && q[2] >= 2 &&
                    q[2] < B + 2))
          continue;
        needed = same.id[hood_index(rel_of(q[0] - 2), rel_of(q[1] - 2), rel_of(q[2] - 2))] < 0;
      }
      if (!needed) continue;
    }
    fill_edge_at<C>(buf, same, coarse, all, bo, po, me, p);
  }
}

// Ghost cell p of the block at `me`: trilinear from the 8 nearest cells of the level
// below (weights 1/4, 3/4 per axis), the 5 fields.
__device__ __forceinline__ void interp_cell(const double* __restrict__ buf, const Hood& coarse,
                                            const int bo[3], const int po[3],
                                            double* __restrict__ me, const int p[3]) {
  int c0[3];
  double w[3][2];
  for (int a = 0; a < 3; ++a) {
    const int g = bo[a] + p[a] - 2;
    const int m = g >> 1;
    const double tt = (g & 1) ? 0.25 : 0.75;
    c0[a] = (g & 1) ? m : m - 1;
    w[a][0] = 1.0 - tt;
    w[a][1] = tt;
  }
  // the 8 coarse cells (field 0 offsets; -1: outside the level below, excluded by
  // the nesting rule), found once for the 5 fields
  long long src[8];
  bool ok[8];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int z = 0; z < 2; ++z) {
        const int i = (x * 2 + y) * 2 + z;
        const int u[3] = {c0[0] + x - po[0], c0[1] + y - po[1], c0[2] + z - po[2]};
        const int r0 = rel_of(u[0]), r1 = rel_of(u[1]), r2 = rel_of(u[2]);
        const int k = hood_index(r0, r1, r2);
        ok[i] = coarse.id[k] >= 0;
        // a missing source reads a valid dummy address and is not summed
        src[i] = ok[i] ? coarse.off[k] + cell_at(0, u[0] - r0 * B + 2, u[1] - r1 * B + 2,
                                                 u[2] - r2 * B + 2)
                       : 0;
      }
  // every load of the 5 fields is issued before the sums, in source order over the
  // sources that exist
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    double v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = buf[src[i] + (long long)f * NCELL];
    double acc = 0.0;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int z = 0; z < 2; ++z) {
          const int i = (x * 2 + y) * 2 + z;
          const double t = acc + ((w[0][x] * w[1][y]) * w[2][z]) * v[i];
          acc = ok[i] ? t : acc;
        }
    me[cell_at(f, p[0], p[1], p[2])] = acc;
  }
}

// The same-level and parent-level neighbourhoods of block d (threads 0..53 load them;
// the caller synchronises).
__device__ __forceinline__ void load_fill_hoods(Hood& same, Hood& coarse,
                                                const BlockDesc* __restrict__ blocks,
                                                const int* __restrict__ lookup, const Levels& lv,
                                                const BlockDesc& d) {
  const int L = d.level;
  const int tid = threadIdx.x;
  if (tid < 54) {
    const bool up = tid >= 27;
    const int k = up ? tid - 27 : tid;
    const int r[3] = {k / 9 - 1, (k / 3) % 3 - 1, k % 3 - 1};
    int id = -1;
    if (!up || L > 0) {
      const int lvl = up ? L - 1 : L;
      const int nb = lv.nb[lvl];
      const int cb[3] = {up ? d.bx >> 1 : d.bx, up ? d.by >> 1 : d.by, up ? d.bz >> 1 : d.bz};
      const int b0 = wrap(cb[0] + r[0], nb), b1 = wrap(cb[1] + r[1], nb),
                b2 = wrap(cb[2] + r[2], nb);
      id = lookup[lv.base[lvl] + (b0 * nb + b1) * nb + b2];
    }
    Hood& h = up ? coarse : same;
    h.id[k] = id;
    h.off[k] = id >= 0 ? blocks[id].off : -1;
  }
}

// Ghosts of every block (workgroups per block: CELL_PARTS for its cells, then one per
// potential component).  A ghost cell takes the value of the block of the same level
// holding it; an edge outside the block's own range [2, 18) takes the value of its
// master copy on the same level (a block keeps the edges it is the master of).  Without
// a block of the same level, cells are interpolated from the level below trilinearly
// (weights 1/4, 3/4 per axis) and edges per axis by the same weights along the cell
// axis and, across, the coarse node (even fine nodes) or the mean of the two coarse
// nodes (odd).  Only values a fill never writes are read, so one launch fills every
// level.
//
// `ids`: the blocks to fill (null: every block).  With `all` (prolongation of new
// blocks) the listed blocks take every cell and edge, interior included, by
// interpolation from the level below.
__global__ void __launch_bounds__(256) k_fill3d(double* __restrict__ buf,
                                                const BlockDesc* __restrict__ blocks,
                                                const int* __restrict__ lookup, Levels lv,
                                                const int* __restrict__ ids, bool all,
                                                bool restricted) {
  __shared__ Hood same, coarse;
  const int self = ids ? ids[blockIdx.x] : (int)blockIdx.x;
  const BlockDesc d = blocks[self];
  const int tid = threadIdx.x;
  load_fill_hoods(same, coarse, blocks, lookup, lv, d);
  __syncthreads();
  const int bo[3] = {d.bx * B, d.by * B, d.bz * B};                    // origin, level L
  const int po[3] = {(d.bx >> 1) * B, (d.by >> 1) * B, (d.bz >> 1) * B};  // parent, L - 1
  double* me = buf + d.off;
  const int part = blockIdx.y;
  // cells: part `part` of them
  const int ncell = part >= CELL_PARTS ? 0 : (all ? NCELL : shell_size(X, X, X));
  const int t_end = (ncell * (part + 1)) / CELL_PARTS;
  for (int t = (ncell * part) / CELL_PARTS + tid; t < t_end; t += blockDim.x) {
    int p[3];
    if (all) {
      p[0] = t / (X * X);
      p[1] = (t / X) % X;
      p[2] = t % X;
    } else {
      shell_at(t, X, X, X, p);
    }
    if (!all) {
      int r[3], lc[3];
      for (int a = 0; a < 3; ++a) {
        r[a] = rel_of(p[a] - 2);
        lc[a] = p[a] - r[a] * B;
      }
      const int k = hood_index(r[0], r[1], r[2]);
      if (same.id[k] >= 0) {
        if (restricted) continue;  // the stage reads it in place
        const double* src = buf + same.off[k];
        const int o = cell_at(0, lc[0], lc[1], lc[2]);
        const int m = cell_at(0, p[0], p[1], p[2]);
        for (int f = 0; f < 5; ++f) me[m + f * NCELL] = src[o + f * NCELL];
        continue;
      }
    }
    interp_cell(buf, coarse, bo, po, me, p);
  }
  // edges of component c = part - CELL_PARTS (axis c is the cell axis, the others
  // node axes)
  if (part == CELL_PARTS) fill_edges<0>(buf, same, coarse, all, restricted, bo, po, me);
  if (part == CELL_PARTS + 1) fill_edges<1>(buf, same, coarse, all, restricted, bo, po, me);
  if (part == CELL_PARTS + 2) fill_edges<2>(buf, same, coarse, all, restricted, bo, po, me);
}

// The ghosts of listed regions of blocks whose level has no block there: a workgroup per
// task {block, region + 32 part, parent}, region (x + 1) * 9 + (y + 1) * 3 + z + 1 of the 26
// around the block (per axis -1: haloed cells 0..1, 0: the block's own 2..B+1, 1:
// B+2..B+3).  It interpolates the region's cells from the level below and sets every
// edge of those cells outside the block's own range as k_fill3d does (the master copy
// on the level, else interpolated), so that it leaves what k_fill3d leaves on the cells
// and edges the stage reads from the block's ghosts: part 0 all of it, part 1 the cells,
// part 2 + c the edges of component c (a face region's work in four workgroups).  An
// edge shared by two regions is written by both tasks with the same value (its sources
// are never written by a fill).  The two neighbourhoods come from `hoods` (per block the
// 27 offsets of the blocks of its level around it, -1: none): the block's and its
// parent's, so a workgroup starts after two dependent loads.
__global__ void __launch_bounds__(256) k_fill3d_regions(double* __restrict__ buf,
                                                        const BlockDesc* __restrict__ blocks,
                                                        const long long* __restrict__ hoods,
                                                        const int4* __restrict__ tasks) {
  __shared__ Hood same, coarse;
  const int4 task = tasks[blockIdx.x];
  const BlockDesc d = blocks[task.x];
  if (threadIdx.x < 54) {
    const bool up = threadIdx.x >= 27;
    const int k = up ? threadIdx.x - 27 : threadIdx.x;
    const int owner = up ? task.z : task.x;
    const long long off = owner >= 0 ? hoods[27 * (size_t)owner + k] : -1;
    Hood& h = up ? coarse : same;
    h.off[k] = off;
    h.id[k] = off >= 0 ? 0 : -1;
  }
  __syncthreads();
  const int bo[3] = {d.bx * B, d.by * B, d.bz * B};
  const int po[3] = {(d.bx >> 1) * B, (d.by >> 1) * B, (d.bz >> 1) * B};
  double* me = buf + d.off;
  const int region = task.y & 31, part = task.y >> 5;  // task.z: the parent block
  const int rel[3] = {region / 9 - 1, (region / 3) % 3 - 1, region % 3 - 1};
  int lo[3], n[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = rel[a] < 0 ? 0 : (rel[a] == 0 ? 2 : B + 2);
    n[a] = rel[a] == 0 ? B : 2;
  }
  const int ncell = part <= 1 ? n[0] * n[1] * n[2] : 0;
  for (int t = threadIdx.x; t < ncell; t += blockDim.x) {
    const int p[3] = {lo[0] + t / (n[1] * n[2]), lo[1] + (t / n[2]) % n[1], lo[2] + t % n[2]};
    interp_cell(buf, coarse, bo, po, me, p);
  }
  // edges of component c: the cell range on axis c, the node range lo .. lo + n on the
  // others (the nodes of the region's cells)
  int m[3][3], ne[3];
  for (int c = 0; c < 3; ++c) {
    for (int a = 0; a < 3; ++a) m[c][a] = n[a] + (a == c ? 0 : 1);
    ne[c] = part == 0 || part == 2 + c ? m[c][0] * m[c][1] * m[c][2] : 0;
  }
  const int nall = ne[0] + ne[1] + ne[2];
  for (int t = threadIdx.x; t < nall; t += blockDim.x) {
    const int c = t < ne[0] ? 0 : (t < ne[0] + ne[1] ? 1 : 2);
    const int r = t - (c == 0 ? 0 : (c == 1 ? ne[0] : ne[0] + ne[1]));
    const int p[3] = {lo[0] + r / (m[c][1] * m[c][2]), lo[1] + (r / m[c][2]) % m[c][1],
                      lo[2] + r % m[c][2]};
    if (p[0] >= 2 && p[0] < B + 2 && p[1] >= 2 && p[1] < B + 2 && p[2] >= 2 && p[2] < B + 2)
      continue;  // the block's own edge
    if (c == 0)
      fill_edge_at<0>(buf, same, coarse, false, bo, po, me, p);
    else if (c == 1)
      fill_edge_at<1>(buf, same, coarse, false, bo, po, me, p);
    else
      fill_edge_at<2>(buf, same, coarse, false, bo, po, me, p);
  }
}

// Covered cells of a refined block := the mean of their 8 children (a workgroup per
// refined block and child; children[9]: the block, then its children with x slowest).
__device__ __forceinline__ void average3d(double* __restrict__ buf,
                                          const BlockDesc* __restrict__ blocks,
                                          const int* __restrict__ row, int oct) {
  double* parent = buf + blocks[row[0]].off;
  // oct: the child (x slowest) whose 8^3 parent cells this group averages
  const double* kid = buf + blocks[row[1 + oct]].off;
  constexpr int H = B / 2;
  for (int t = threadIdx.x; t < H * H * H; t += blockDim.x) {
    const int li = t / (H * H), lj = (t / H) % H, lk = t % H;
    const int I = (oct >> 2) * H + li, J = ((oct >> 1) & 1) * H + lj, K = (oct & 1) * H + lk;
    const int fi = 2 * li + 2, fj = 2 * lj + 2, fk = 2 * lk + 2;
    for (int f = 0; f < 5; ++f) {
      const double s0 = (kid[cell_at(f, fi, fj, fk)] + kid[cell_at(f, fi, fj, fk + 1)]) +
                        (kid[cell_at(f, fi, fj + 1, fk)] + kid[cell_at(f, fi, fj + 1, fk + 1)]);
      const double s1 = (kid[cell_at(f, fi + 1, fj, fk)] + kid[cell_at(f, fi + 1, fj, fk + 1)]) +
                        (kid[cell_at(f, fi + 1, fj + 1, fk)] + kid[cell_at(f, fi + 1, fj + 1, fk + 1)]);
      parent[cell_at(f, I + 2, J + 2, K + 2)] = (s0 + s1) * 0.125;
    }
  }
}

// Coarse edges of a refined block := the mean of the two fine edges along them
// (mhd3d.py:31-36), written to every block of the coarse level holding the edge.  A
// workgroup per refined block, component c and part of its edges; `hood`: the offsets
// of the 27 blocks of the coarse level around the block (-1: none), `kids`: those of its
// 8 children (x slowest), both in LDS.
__device__ __forceinline__ void restrict3d(double* __restrict__ buf, const long long* hood,
                                           const long long* kids, int c, int part) {
  const int e0 = c == 0 ? B : B + 1, e1 = c == 1 ? B : B + 1, e2 = c == 2 ? B : B + 1;
  const int ne = e0 * e1 * e2, t_end = (ne * (part + 1)) / RESTRICT_PARTS;
  for (int t = (ne * part) / RESTRICT_PARTS + threadIdx.x; t < t_end; t += blockDim.x) {
    const int q[3] = {t / (e1 * e2), (t / e2) % e1, t % e2};  // 0..B (nodes) / B-1 (cells)
    // the two fine edges: fine coordinates within the block's footprint (0..2B)
    double v[2];
    for (int s = 0; s < 2; ++s) {
      int oct[3], loc[3];
      for (int a = 0; a < 3; ++a) {
        const int fcoord = 2 * q[a] + (a == c ? s : 0);
        oct[a] = fcoord >= 2 * B ? 1 : fcoord / B;
        loc[a] = fcoord - B * oct[a] + 2;
      }
      v[s] = buf[kids[(oct[0] * 2 + oct[1]) * 2 + oct[2]] + edge_at(c, loc[0], loc[1], loc[2])];
    }
    const double value = 0.5 * (v[0] + v[1]);
    // every coarse block holding edge q of the block: the block containing it (q = B
    // on a node axis: the next block, local index 2) and, for a node coordinate on a
    // block boundary, the block below it (local index 2 + B)
    for (int s0 = 0; s0 < 2; ++s0)
      for (int s1 = 0; s1 < 2; ++s1)
        for (int s2 = 0; s2 < 2; ++s2) {
          const int sh[3] = {s0, s1, s2};
          int rel[3], loc[3];
          bool ok = true;
          for (int a = 0; a < 3; ++a) {
            if (sh[a] == 0) {
              rel[a] = q[a] / B;
              loc[a] = q[a] % B + 2;
            } else {  // the copy on the upper boundary of the block below
              ok = ok && a != c && q[a] % B == 0;
              rel[a] = q[a] / B - 1;
              loc[a] = B + 2;
            }
          }
          if (!ok) continue;
          const long long off = hood[(rel[0] + 1) * 9 + (rel[1] + 1) * 3 + rel[2] + 1];
          if (off >= 0) buf[off + edge_at(c, loc[0], loc[1], loc[2])] = value;
        }
  }
}

// One level pair's coarsening in one launch: per refined block (blockIdx.x), 8
// workgroups average its cells (one per child), then 3 x RESTRICT_PARTS restrict its
// edges; the two write disjoint data (cells, edges) from disjoint data.  first = 8
// (a grid of 3 x RESTRICT_PARTS) restricts the edges only.
__global__ void __launch_bounds__(256) k_coarsen3d(double* __restrict__ buf,
                                                   const BlockDesc* __restrict__ blocks,
                                                   const int* __restrict__ children,
                                                   const int* __restrict__ lookup, Levels lv,
                                                   int first) {
  __shared__ long long hood[27], kids[8];
  const int* row = children + 9 * blockIdx.x;
  const int y = blockIdx.y + first;
  if (y < 8) {
    average3d(buf, blocks, row, y);
    return;
  }
  load_hood(hood, blocks, lookup, lv, blocks[row[0]]);
  if (threadIdx.x >= 32 && threadIdx.x < 40) kids[threadIdx.x - 32] = blocks[row[threadIdx.x - 31]].off;
  __syncthreads();
  restrict3d(buf, hood, kids, (y - 8) / RESTRICT_PARTS, (y - 8) % RESTRICT_PARTS);
}



}  // namespace

namespace {

// The flux register and refluxing (FluxRegister.update with scale 0.5 dt and
// FluxRegister.reflux as MHD3D.step calls them, mhd3d.py:650-655,745-750): per
// coarse-fine block face, the 8 x 8 coarse cells across it trade the flux of their face
// on the fine side for the mean of the four fine fluxes.  Both are the fluxes a stage
// solved on those block faces (recorded by the stage with the blocks' 6 face slots
// `slots`: stage 1 in rec1, stage 2 in rec): d = sum over the cell's faces of dt/h (F_c -
// mean) or dt/h (mean - F_c).  After stage 2 and its Heun average the cell gains
// reg + 0.5 d2, reg = 0.5 d1 being what the register holds after stage 1: both stages'
// refluxing in one launch.  An item is one such face (pad = 2 axis + side;
// side 0: the coarse cells at the fine block's low face, their +axis face, the coarse
// block's high face; side 1: at the fine block's high face, the coarse block's low face).
// `cells` lists, per coarse cell, its n (1 to 3) faces as (item, lane) pairs in
// increasing direction order, the lane u * 8 + v naming the cell inside the item's 8 x 8
// face; a thread takes one cell and sums its faces' corrections in that order, with no
// atomics on the state (deterministic).
__global__ void __launch_bounds__(256) k_reflux3d(const BlockDesc* __restrict__ blocks,
                                                  const RefluxItem* __restrict__ items,
                                                  const RefluxCell* __restrict__ cells, int ncells,
                                                  const int* __restrict__ slots,
                                                  const double* __restrict__ rec1,
                                                  const double* __restrict__ rec,
                                                  const double* __restrict__ dtp,
                                                  double* __restrict__ dst) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells) return;
  const RefluxCell cl = cells[t];
  const double dt = *dtp;
  // acc1, acc: the faces' corrections from the fluxes stage 1 (rec1) and stage 2 (rec)
  // recorded, summed alike
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, acc1[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int c[3] = {0, 0, 0};
  BlockDesc cd{};
  for (int f = 0; f < cl.n; ++f) {
    const RefluxItem item = items[cl.item[f]];
    const int lc = cl.lane[f];
    const int axis = item.pad >> 1, side = item.pad & 1;
    const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
    const int u = lc >> 3, v = lc & 7;
    // the coarse cell (haloed local) and its face on the fine side
    c[0] = item.ci + 2;
    c[1] = item.cj + 2;
    c[2] = item.ck + 2;
    c[a1] += u;
    c[a2] += v;
    cd = blocks[item.coarse];
    const int cs = slots[6 * item.coarse + 2 * axis + (1 - side)];
    const int fs = slots[6 * item.fine + 2 * axis + side];
    const size_t oc = ((size_t)cs * (B * B) + (c[a1] - 2) * B + (c[a2] - 2)) * 5;
    size_t of[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      of[k] = ((size_t)fs * (B * B) + (2 * u + (k >> 1)) * B + 2 * v + (k & 1)) * 5;
    const double dtd = axis == 0 ? dt / cd.dx : (axis == 1 ? dt / cd.dy : dt / cd.dz);
    auto add = [&](const double* r, double* sum) {
#pragma unroll
      for (int fld = 0; fld < 5; ++fld) {
        const double fm =
            ((r[of[0] + fld] + r[of[1] + fld]) + (r[of[2] + fld] + r[of[3] + fld])) * 0.25;
        const double d = side == 0 ? dtd * (r[oc + fld] - fm) : dtd * (fm - r[oc + fld]);
        sum[fld] = sum[fld] + d;
      }
    };
    add(rec, acc);
    add(rec1, acc1);
  }
  // the cell gains the register (stage 1's half correction) and stage 2's half
  double* out = dst + cd.off;
  for (int fld = 0; fld < 5; ++fld) {
    const int o = cell_at(fld, c[0], c[1], c[2]);
    const double reg = 0.5 * acc1[fld];
    out[o] = out[o] + (reg + 0.5 * acc[fld]);
  }
}

}  // namespace

extern "C" {

// Hidden: the stage arguments of a block (shared with the mixed unit).
__attribute__((visibility("hidden"))) void wb_amr3d_base_args(
    void* args, const double* in, double* out, double b0x, double b0y, double b0z, double gamma,
    int limiter, const double* dt, int heun, int32_t* bad) {
  StageArgs& a = *(StageArgs*)args;
  a.in = in;
  a.out = out;
  a.dt = dt;
  a.bad = (int*)bad;
  a.g = Grid3{X, X, X, B, B, B};
  a.dx = a.dy = a.dz = 1.0;
  a.b0x = b0x;
  a.b0y = b0y;
  a.b0z = b0z;
  a.gamma = gamma;
  a.limiter = limiter;
  a.heun = heun;
  a.nxs = B;
  a.wrap = 0;
  a.off_ax = OFF_A;
  a.off_ay = OFF_A + NEDGE;
  a.off_az = OFF_A + 2 * NEDGE;
}

int wb_amr3d_fill(double* buf, const void* blocks, int nblocks, const int32_t* lookup,
                  const void* levels, void* stream) {
  const Levels lv = *(const Levels*)levels;
  if (nblocks < 1 || !levels_ok(lv)) return WB_EINVAL;
  hipLaunchKernelGGL(k_fill3d, dim3(nblocks, CELL_PARTS + 3), dim3(256), 0, (hipStream_t)stream, buf,
                     (const BlockDesc*)blocks, (const int*)lookup, lv, (const int*)nullptr, false,
                     false);
  return wb_last_error();
}

int wb_amr3d_fill_list(double* buf, const void* blocks, const int32_t* ids, int nids,
                       int restricted, const int32_t* lookup, const void* levels, void* stream) {
  const Levels lv = *(const Levels*)levels;
  if (nids < 0 || !ids || !levels_ok(lv) || restricted < 0 || restricted > 1) return WB_EINVAL;
  if (nids == 0) return 0;
  hipLaunchKernelGGL(k_fill3d, dim3(nids, CELL_PARTS + 3), dim3(256), 0, (hipStream_t)stream, buf,
                     (const BlockDesc*)blocks, (const int*)lookup, lv, (const int*)ids, false,
                     restricted != 0);
  return wb_last_error();
}

int wb_amr3d_prolong(double* buf, const void* blocks, const int32_t* ids, int nids,
                     const int32_t* lookup, const void* levels, void* stream) {
  const Levels lv = *(const Levels*)levels;
  if (nids < 0 || !levels_ok(lv)) return WB_EINVAL;
  if (nids == 0) return 0;
  hipLaunchKernelGGL(k_fill3d, dim3(nids, CELL_PARTS + 3), dim3(256), 0, (hipStream_t)stream, buf,
                     (const BlockDesc*)blocks, (const int*)lookup, lv, (const int*)ids, true,
                     false);
  return wb_last_error();
}

int wb_amr3d_fill_regions(double* buf, const void* blocks, const long long* hoods,
                          const int32_t* tasks, int ntasks, void* stream) {
  if (ntasks < 0) return WB_EINVAL;
  if (ntasks == 0) return 0;
  hipLaunchKernelGGL(k_fill3d_regions, dim3((unsigned)ntasks), dim3(256), 0, (hipStream_t)stream,
                     buf, (const BlockDesc*)blocks, hoods, (const int4*)tasks);
  return wb_last_error();
}

int wb_amr3d_coarsen(double* buf, const void* blocks, const int32_t* children, int nrefined,
                     int cells, const int32_t* lookup, const void* levels, void* stream) {
  const Levels lv = *(const Levels*)levels;
  if (nrefined < 0 || !levels_ok(lv) || cells < 0 || cells > 1) return WB_EINVAL;
  if (nrefined == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const int first = cells ? 0 : 8;
  hipLaunchKernelGGL(k_coarsen3d, dim3(nrefined, 8 + 3 * RESTRICT_PARTS - first), dim3(256), 0, s,
                     buf, (const BlockDesc*)blocks, (const int*)children, (const int*)lookup, lv,
                     first);
  return wb_last_error();
}

int wb_amr3d_rate(const double* buf, const void* blocks, int nblocks, double b0x, double b0y,
                  double b0z, double gamma, double* out, void* stream) {
  if (nblocks < 1) return WB_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  StageArgs a;
  wb_amr3d_base_args(&a, buf, const_cast<double*>(buf), b0x, b0y, b0z, gamma, 0, nullptr, 0,
                     nullptr);
  unsigned long long* keys = (unsigned long long*)out;
  hipLaunchKernelGGL(k_rate_init, dim3(1), dim3(64), 0, s, keys);
  hipLaunchKernelGGL(k_rate3d_blocks<false>, dim3((unsigned)nblocks), dim3(256 * RATE_HALVES), 0, s, a,
                     (const BlockDesc*)blocks, keys);
  hipLaunchKernelGGL(k_rate_finish, dim3(1), dim3(64), 0, s, keys, out);
  return wb_last_error();
}

int wb_amr3d_stage(const double* in, double* out, const void* blocks, const int32_t* ids,
                   int nids, int direct, const int32_t* lookup, const void* levels, double b0x,
                   double b0y, double b0z, double gamma, int limiter, const double* dt, int heun,
                   int32_t* bad, const int32_t* rec_slots, double* rec, void* stream) {
  if (nids < 1 || limiter < 0 || limiter > 3 || direct < 0 || direct > 1) return WB_EINVAL;
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
  const dim3 grid(((nids + 7) / 8) * 8), tile(Tile<B, B>::THREADS);
  hipStream_t s = (hipStream_t)stream;
  if (direct)
    hipLaunchKernelGGL((k_stage3d_blocks<double, 1, -1, true>), grid, tile, 0, s, a,
                       (const BlockDesc*)blocks, (const int*)ids, nids, (const int*)lookup, lv);
  else
    hipLaunchKernelGGL((k_stage3d_blocks<double, 1, -1, false>), grid, tile, 0, s, a,
                       (const BlockDesc*)blocks, (const int*)ids, nids, (const int*)lookup, lv);
  return wb_last_error();
}

int wb_amr3d_reflux_both(double* dst, const void* blocks, const void* items, const void* cells,
                         int ncells, const int32_t* rec_slots, const double* rec1,
                         const double* rec2, const double* dt, void* stream) {
  if (ncells < 0 ||
      (ncells > 0 && (rec_slots == nullptr || rec1 == nullptr || rec2 == nullptr)))
    return WB_EINVAL;
  if (ncells == 0) return 0;
  hipLaunchKernelGGL(k_reflux3d, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, (const BlockDesc*)blocks, (const RefluxItem*)items,
                     (const RefluxCell*)cells, ncells, (const int*)rec_slots, rec1, rec2, dt,
                     dst);
  return wb_last_error();
}

}  // extern "C"
