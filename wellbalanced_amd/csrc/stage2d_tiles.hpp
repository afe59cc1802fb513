// The 2D stage over a refined hierarchy's patches in square tiles: one workgroup of 256
// threads per 16 x 16 cells of a patch, all tiles of all patches in one launch.
//
// The marching stage (stage2d.hpp) walks a strip of columns along x and keeps the
// x-sweep in registers; on patches of 16..128 cells its strips are short and narrow, and
// a hierarchy of a few hundred thousand cells gives too few waves to hide the latency of
// each plane's chain of loads, barriers and face solves.  Here a workgroup does its tile
// in three passes with no march:
//   1. the primitive variables of the 20 x 20 cells around the tile (2-cell halo, the
//      ghosts filled before the launch) into LDS, with each cell's low-face fields (the
//      B_n of the faces the tile solves) and the cell E_z of the 18 x 18 cells the CT
//      reads (`mhd_primitive_pyccel.py:51-86`);
//   2. the 17 x 18 x-faces and 18 x 17 y-faces (`mhd_sweep_pyccel.py:254-319`): the x-
//      and y-face tasks are padded to 320 each, so every wave solves faces of one kind;
//   3. per thread one cell and its low-corner node: the CT-contact E_z of the node
//      (`ct_contact_pyccel.py:35-61`), A_z -= dt E_z, the conservative update of the cell,
//      the Heun average; the nodes on a patch's high faces in a short extra pass.
// Every flux, E_z and update is the marching stage's expression on the same values, so
// both give the same bits (the FP64 instance is the parity path).  NI adds the resistive
// and heat-conduction face terms (FP64) as the marching stage does: the first pass also
// keeps T = p / rho and the cell field of the loaded cells and J_z at the 19 x 19 nodes
// around the tile in LDS.
#pragma once

#include "mhd_device.hpp"
#include "mhd_fast.hpp"
#include "stage2d.hpp"
#include "wbmhd_common.hpp"

namespace wb {
namespace tiles2d {

constexpr int T = 16;               // cells per tile side
constexpr int H = T + 4;            // cells loaded per side (2-cell halo)
constexpr int NC = H * H;           // loaded cells
constexpr int E = T + 2;            // cells with a cell E_z per side ([-1, T])
constexpr int NXF = (T + 1) * (T + 2);  // x-faces: x-node [0, T] x cell [-1, T]
constexpr int NYF = (T + 2) * (T + 1);  // y-faces: cell [-1, T] x y-node [0, T]
constexpr int PAD = 320;            // face tasks of one kind, a whole number of waves
constexpr int THREADS = T * T;
constexpr int N = T + 3;            // NI: nodes with a J_z per side ([-1, T + 1])
static_assert(NXF <= PAD && NYF <= PAD && PAD % 64 == 0, "face tasks fit whole waves");

template <typename P, bool NI = false>
struct Shared {
  P prim[8][NC];    // rho, u, v, w, p, B_x, B_y, B_z of cell (i, j), local [-2, T+1]
  P bxf[NC];        // B_n of the cell's low x-face
  P byf[NC];        // B_n of the cell's low y-face
  double ecc[E * E];  // cell E_z, local [-1, T]
  P fx[7][NXF];     // x-face (x-node i, cell j): index i * (T + 2) + j + 1
  P fy[7][NYF];     // y-face (cell i, y-node j): index (i + 1) * (T + 1) + j
  // NI: T = p / rho, B_x, B_y, B_z of the loaded cells; J_z at node (i, j), local [-1, T+1]
  double cell[NI ? 4 : 1][NI ? NC : 1];
  double jz[NI ? N * N : 1];
};

__device__ __forceinline__ int cell_index(int i, int j) { return (i + 2) * H + (j + 2); }
__device__ __forceinline__ int xface_index(int i, int j) { return i * (T + 2) + j + 1; }
__device__ __forceinline__ int yface_index(int i, int j) { return (i + 1) * (T + 1) + j; }
__device__ __forceinline__ int node_index(int i, int j) { return (i + 1) * N + (j + 1); }

// Tile (tx, ty) of the patch whose stage arguments are `a` (in, out, nx, ny, X, Y, dx,
// dy, walls set for that patch).  `recs` (count, then count pairs (code, slot); null:
// none): faces whose fluxes the refluxing reads, code = kind | index << 1 (kind 0: the
// x-face fx index, 1: the y-face fy index); each is stored as the 6 fluxes of the cell
// fields (x-rows 0, 1, 2, 3, 4, 6 or y-rows 0, 2, 1, 3, 4, 6), widened to FP64, at
// rec + 6 slot.
template <typename P, int LIM, bool NI = false>
__device__ __forceinline__ void tile_body(const Stage2Args& a, int tx, int ty, Shared<P, NI>& sh,
                                          const int* recs = nullptr, double* rec = nullptr) {
  const int tid = threadIdx.x;
  const int X = a.X, Y = a.Y, YN = Y + 1;
  const int i0 = 2 + tx * T, j0 = 2 + ty * T;  // haloed index of the tile's first cell
  const int txv = min(T, a.nx + 2 - i0), tyv = min(T, a.ny + 2 - j0);
  const bool last_x = i0 + T >= a.nx + 2, last_y = j0 + T >= a.ny + 2;
  const size_t PL = (size_t)X * Y;
  const double* az = a.in + 6 * PL;
  auto AZ = [&](int i, int j) -> double { return az[(size_t)i * YN + j]; };

  // the stored values this thread's update starts from -- the 6 fields of cell (ci, cj)
  // and the potential of its low-corner node, in `in` and (Heun average) in `out` --
  // loaded before the passes so that their latency overlaps them; no other workgroup
  // writes them
  const int ci = tid / T, cj = tid % T;
  const bool own_x = ci < txv || (ci == txv && last_x), own_y = cj < tyv || (cj == tyv && last_y);
  const bool has_cell = ci < txv && cj < tyv, has_node = own_x && own_y;
  double cin[6], cout[6], nin = 0.0, nout = 0.0;
  {
    const size_t c = has_cell ? (size_t)(i0 + ci) * Y + (j0 + cj) : 0;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      cin[r] = has_cell ? a.in[r * PL + c] : 0.0;
      cout[r] = has_cell && a.heun ? a.out[r * PL + c] : 0.0;
    }
    const size_t o = has_node ? 6 * PL + (size_t)(i0 + ci) * YN + (j0 + cj) : 0;
    nin = has_node ? a.in[o] : 0.0;
    nout = has_node && a.heun ? a.out[o] : 0.0;
  }

  // 1. primitive variables of the loaded cells (mhd_primitive_pyccel.py:51-86)
  for (int k = tid; k < NC; k += THREADS) {
    const int li = k / H - 2, lj = k % H - 2;
    const int i = min(max(i0 + li, 0), X - 1), j = min(max(j0 + lj, 0), Y - 1);
    const size_t c = (size_t)i * Y + j;
    const double bx0 = (AZ(i, j + 1) - AZ(i, j)) / a.dy + a.b0x;
    const double bx1 = (AZ(i + 1, j + 1) - AZ(i + 1, j)) / a.dy + a.b0x;
    const double by0 = -(AZ(i + 1, j) - AZ(i, j)) / a.dx + a.b0y;
    const double by1 = -(AZ(i + 1, j + 1) - AZ(i, j + 1)) / a.dx + a.b0y;
    const double r = a.in[c];
    const double u = a.in[PL + c] / r, v = a.in[2 * PL + c] / r, w = a.in[3 * PL + c] / r;
    const double bx = 0.5 * (bx0 + bx1);
    const double by = 0.5 * (by0 + by1);
    const double b = a.in[5 * PL + c];
    const double p = (a.gamma - 1.0) * (a.in[4 * PL + c] - 0.5 * r * (u * u + v * v + w * w) -
                                        0.5 * (bx * bx + by * by + b * b));
    sh.prim[0][k] = P(r);
    sh.prim[1][k] = P(u);
    sh.prim[2][k] = P(v);
    sh.prim[3][k] = P(w);
    sh.prim[4][k] = P(p);
    sh.prim[5][k] = P(bx);
    sh.prim[6][k] = P(by);
    sh.prim[7][k] = P(b);
    sh.bxf[k] = P(bx0);
    sh.byf[k] = P(by0);
    if (li >= -1 && li <= T && lj >= -1 && lj <= T)
      sh.ecc[(li + 1) * E + (lj + 1)] = bx * v - by * u;
    if constexpr (NI) {
      sh.cell[0][k] = p / r;
      sh.cell[1][k] = bx;
      sh.cell[2][k] = by;
      sh.cell[3][k] = b;
    }
    if (li >= 0 && li < txv && lj >= 0 && lj < tyv) {
      const bool is_bad = !(r > 0.0 && p > 0.0) ||
                          !(finite(r) && finite(u) && finite(v) && finite(w) && finite(p) &&
                            finite(bx) && finite(by) && finite(b));
      if (is_bad) atomicAdd(a.bad, 1);
    }
  }
  if constexpr (NI) {  // J_z at the nodes around the tile, inside the haloed patch
    for (int k = tid; k < N * N; k += THREADS) {
      const int I = min(max(i0 + k / N - 1, 1), X - 1), J = min(max(j0 + k % N - 1, 1), Y - 1);
      sh.jz[k] = node_current_az(az, YN, I, J, a.dx, a.dy, a.b0x, a.b0y);
    }
  }
  __syncthreads();

  // 2. face fluxes; tasks [0, PAD) are x-faces, [PAD, 2 PAD) y-faces (wave-uniform kinds)
  const P gam = P(a.gamma), inv_gm1 = P(1.0 / (a.gamma - 1.0));
  for (int k = tid; k < 2 * PAD; k += THREADS) {
    if (k < PAD) {
      if (k < NXF) {
        const int i = k / (T + 2), j = k % (T + 2) - 1;
        constexpr int comp[7] = {0, 1, 2, 3, 4, 6, 7};
        P q[7][4];
#pragma unroll
        for (int c = 0; c < 7; ++c)
#pragma unroll
          for (int s = 0; s < 4; ++s) q[c][s] = sh.prim[comp[c]][cell_index(i - 2 + s, j)];
        P fl[7];
        face_solve<P, LIM>(q, a.limiter, sh.bxf[cell_index(i, j)], gam, inv_gm1, fl);
        if constexpr (NI) {  // mhd.py:868-871, cells L = (i - 1, j) and R = (i, j)
          const int l = cell_index(i - 1, j), c = cell_index(i, j);
          const int fi = min(i0 + i, X - 1), jj = min(max(j0 + j, 0), Y - 1);
          const double bnd = (AZ(fi, jj + 1) - AZ(fi, jj)) / a.dy + a.b0x;
          double e4 = double(fl[4]), b6 = double(fl[6]);
          if (a.diff.conduction())
            conductive_face(bnd, sh.cell[2][l], sh.cell[2][c], sh.cell[3][l], sh.cell[3][c],
                            sh.cell[0][l], sh.cell[0][c],
                            limited_gradient(sh.cell[0][l - 1], sh.cell[0][l], sh.cell[0][l + 1], a.dy),
                            limited_gradient(sh.cell[0][c - 1], sh.cell[0][c], sh.cell[0][c + 1], a.dy),
                            a.dx, a.diff, e4);
          if (a.diff.eta != 0.0)
            resistive_xface(sh.cell[3][l], sh.cell[3][c], sh.cell[2][l], sh.cell[2][c],
                            sh.jz[node_index(i, j)], sh.jz[node_index(i, j + 1)], a.dx,
                            a.diff.eta, e4, b6);
          fl[4] = P(e4);
          fl[6] = P(b6);
        }
#pragma unroll
        for (int r = 0; r < 7; ++r) sh.fx[r][k] = fl[r];
      }
    } else {
      const int m = k - PAD;
      if (m < NYF) {
        const int i = m / (T + 1) - 1, j = m % (T + 1);
        constexpr int comp[7] = {0, 2, 1, 3, 4, 5, 7};
        P q[7][4];
#pragma unroll
        for (int c = 0; c < 7; ++c)
#pragma unroll
          for (int s = 0; s < 4; ++s) q[c][s] = sh.prim[comp[c]][cell_index(i, j - 2 + s)];
        P fl[7];
        face_solve<P, LIM>(q, a.limiter, sh.byf[cell_index(i, j)], gam, inv_gm1, fl);
        if constexpr (NI) {  // cells L = (i, j - 1) and R = (i, j)
          const int l = cell_index(i, j - 1), c = cell_index(i, j);
          const int ii = min(max(i0 + i, 0), X - 1), jj = min(max(j0 + j, 0), Y - 1);
          const double bnd = -(AZ(ii + 1, jj) - AZ(ii, jj)) / a.dx + a.b0y;
          double e4 = double(fl[4]), b6 = double(fl[6]);
          if (a.diff.conduction())
            conductive_face(bnd, sh.cell[1][l], sh.cell[1][c], sh.cell[3][l], sh.cell[3][c],
                            sh.cell[0][l], sh.cell[0][c],
                            limited_gradient(sh.cell[0][l - H], sh.cell[0][l], sh.cell[0][l + H], a.dx),
                            limited_gradient(sh.cell[0][c - H], sh.cell[0][c], sh.cell[0][c + H], a.dx),
                            a.dy, a.diff, e4);
          if (a.diff.eta != 0.0)
            resistive_yface(sh.cell[3][l], sh.cell[3][c], sh.cell[1][l], sh.cell[1][c],
                            sh.jz[node_index(i, j)], sh.jz[node_index(i + 1, j)], a.dy,
                            a.diff.eta, e4, b6);
          fl[4] = P(e4);
          fl[6] = P(b6);
        }
#pragma unroll
        for (int r = 0; r < 7; ++r) sh.fy[r][m] = fl[r];
      }
    }
  }
  __syncthreads();

  // the fluxes the refluxing reads (the values the update below applies)
  if (recs != nullptr) {
    const int count = recs[0];
    for (int k = tid; k < count; k += THREADS) {
      const int code = recs[1 + 2 * k], slot = recs[2 + 2 * k];
      const int idx = code >> 1;
      double* o = rec + 6 * (size_t)slot;
      if ((code & 1) == 0) {
        constexpr int XR[6] = {0, 1, 2, 3, 4, 6};
#pragma unroll
        for (int r = 0; r < 6; ++r) o[r] = double(sh.fx[XR[r]][idx]);
      } else {
        constexpr int YR[6] = {0, 2, 1, 3, 4, 6};
#pragma unroll
        for (int r = 0; r < 6; ++r) o[r] = double(sh.fy[YR[r]][idx]);
      }
    }
  }

  // 3. the CT at node (i, j) and the update of cell (i, j)
  const double dt = *a.dt;
  const double dtdx = dt / a.dx, dtdy = dt / a.dy;
  const bool xlo = a.walls & 1, xhi = a.walls & 2, ylo = a.walls & 4, yhi = a.walls & 8;
  // node (i, j); `pre`: its stored values are vin, vout (loaded above)
  auto node = [&](int i, int j, bool pre, double vin, double vout) {
    const int f = i0 + i, g = j0 + j;
    double e = ct_node(double(sh.fx[5][xface_index(i, j - 1)]), double(sh.fx[5][xface_index(i, j)]),
                       double(sh.fy[5][yface_index(i - 1, j)]), double(sh.fy[5][yface_index(i, j)]),
                       double(sh.fx[0][xface_index(i, j - 1)]), double(sh.fx[0][xface_index(i, j)]),
                       double(sh.fy[0][yface_index(i - 1, j)]), double(sh.fy[0][yface_index(i, j)]),
                       sh.ecc[i * E + j], sh.ecc[(i + 1) * E + j], sh.ecc[i * E + j + 1],
                       sh.ecc[(i + 1) * E + j + 1]);
    if constexpr (NI) {
      if (a.diff.eta != 0.0) e = e + a.diff.eta * sh.jz[node_index(i, j)];
    }
    if ((xlo && f == 2) || (xhi && f == a.nx + 2) || (ylo && g == 2) || (yhi && g == a.ny + 2))
      e = 0.0;
    const size_t o = 6 * PL + (size_t)f * YN + g;
    double v = (pre ? vin : a.in[o]) - dt * e;
    if (a.heun) v = (v + (pre ? vout : a.out[o])) * 0.5;
    a.out[o] = v;
  };
  if (has_node) node(ci, cj, true, nin, nout);
  if (has_cell) {
    const int f = i0 + ci, g = j0 + cj;
    constexpr int XR[6] = {0, 1, 2, 3, 4, 6};  // x-sweep rows of the 6 fields (mhd.py:134)
    constexpr int YR[6] = {0, 2, 1, 3, 4, 6};
    const bool wxl = (xlo && f == 2), wxh = (xhi && f + 1 == a.nx + 2);
    const bool wyl = ylo && g == 2, wyh = yhi && g + 1 == a.ny + 2;
    const size_t c = (size_t)f * Y + g;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      const int xr = XR[r], yr = YR[r];
      const bool xz = xr != 1 && xr != 5, yz = yr != 1 && yr != 5;
      const P flo = (wxl && xz) ? P(0) : sh.fx[xr][xface_index(ci, cj)];
      const P fhi = (wxh && xz) ? P(0) : sh.fx[xr][xface_index(ci + 1, cj)];
      const double dfx = double(fhi) - double(flo);
      const double hi = (wyh && yz) ? 0.0 : double(sh.fy[yr][yface_index(ci, cj + 1)]);
      const double lo = (wyl && yz) ? 0.0 : double(sh.fy[yr][yface_index(ci, cj)]);
      const double dyy = dtdy * (hi - lo);
      const size_t o = r * PL + c;
      double v = cin[r] - (dtdx * dfx + dyy);
      if (a.heun) v = (v + cout[r]) * 0.5;
      a.out[o] = v;
    }
  }
  // the nodes on the tile's high faces (i == T or j == T) where they are the patch's
  if ((last_x && txv == T) || (last_y && tyv == T)) {
    for (int k = tid; k < 2 * T + 1; k += THREADS) {
      const int i = k <= T ? T : k - T - 1, j = k <= T ? k : T;
      const bool ox = i < txv || (i == txv && last_x), oy = j < tyv || (j == tyv && last_y);
      if (ox && oy) node(i, j, false, 0.0, 0.0);
    }
  }
}

}  // namespace tiles2d

// One workgroup per item (patch, tile along y, tile along x: WorkItem's patch, tile, seg),
// or, without a patch table (`patches` null), per tile of the one grid `base` describes
// (tiles along y fastest).  Workgroup w runs on XCD w % 8; each XCD takes a contiguous
// run of the items, so the tiles that share halo cells share an L2.
// rec_first (null: no record): per item the start of its record list in rec_lists.
template <typename P, int LIM, bool NI>
__global__ void __launch_bounds__(tiles2d::THREADS)
k_stage2d_tiles(Stage2Args base, const PatchDesc* __restrict__ patches,
                const WorkItem* __restrict__ items, int nitems, double* __restrict__ out_base,
                const int* __restrict__ rec_first, const int* __restrict__ rec_lists,
                double* __restrict__ rec) {
  __shared__ tiles2d::Shared<P, NI> sh;
  const int n = gridDim.x, xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
  const int per = n >> 3, rem = n & 7;
  const int it = xcd * per + min(xcd, rem) + k;
  if (it >= nitems) return;
  if (patches == nullptr) {
    const int tiles_y = (base.ny + tiles2d::T - 1) / tiles2d::T;
    tiles2d::tile_body<P, LIM, NI>(base, it / tiles_y, it % tiles_y, sh);
    return;
  }
  const WorkItem w = items[it];
  const PatchDesc pd = patches[w.patch];
  Stage2Args a = base;
  a.in = base.in + pd.off;
  a.out = out_base + pd.off;
  a.nx = pd.nx;
  a.ny = pd.ny;
  a.X = pd.nx + 4;
  a.Y = pd.ny + 4;
  a.dx = pd.dx;
  a.dy = pd.dy;
  a.walls = pd.walls;
  const int* recs = rec_first ? rec_lists + rec_first[it] : nullptr;
  tiles2d::tile_body<P, LIM, NI>(a, w.seg, w.tile, sh, recs, rec);
}

// patches == nullptr: every tile of the grid `a` (nitems is then the tile count); the
// non-ideal instance where a.diff has a resistivity or a conductivity
template <typename P, int LIM>
inline int launch_stage2d_tiles(const Stage2Args& a, const void* patches, const void* items,
                                int nitems, double* out, hipStream_t s,
                                const int32_t* rec_first = nullptr,
                                const int32_t* rec_lists = nullptr, double* rec = nullptr) {
  if (a.diff.conduction() || a.diff.eta != 0.0)
    hipLaunchKernelGGL((k_stage2d_tiles<P, LIM, true>), dim3((unsigned)nitems),
                       dim3(tiles2d::THREADS), 0, s, a, (const PatchDesc*)patches,
                       (const WorkItem*)items, nitems, out, (const int*)rec_first,
                       (const int*)rec_lists, rec);
  else
    hipLaunchKernelGGL((k_stage2d_tiles<P, LIM, false>), dim3((unsigned)nitems),
                       dim3(tiles2d::THREADS), 0, s, a, (const PatchDesc*)patches,
                       (const WorkItem*)items, nitems, out, (const int*)rec_first,
                       (const int*)rec_lists, rec);
  return wb_last_error();
}

// The flux register and refluxing from the fluxes the tiled stage recorded (the
// arithmetic of k_reflux2d on the same values): per group (coarse cell) rows of 8 int32
// (coarse patch, cell a, cell b, axis, side, slot of the coarse face, slots of the two
// fine faces), from the fluxes of stage 1 (rec1) and stage 2 (rec) after the Heun average:
// cell += reg + 0.5 d2, reg = 0.5 d1 being what the register holds after stage 1 (both
// stages' refluxing in one launch; the product by 0.5 is exact, so the sum rounds as
// with a register kept in memory).  Internal linkage: each translation unit compiles it
// with its own contraction rule, as it does the stage.
static __global__ void __launch_bounds__(64)
k_reflux2d_recorded(double* __restrict__ dst, const PatchDesc* __restrict__ patches,
                    const int* __restrict__ rows, const int* __restrict__ groups, int ngroups,
                    const double* __restrict__ rec1, const double* __restrict__ rec,
                    const double* __restrict__ dtp) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  const double dt = *dtp;
  const int first = groups[2 * g], count = groups[2 * g + 1];
  double delta[6] = {0, 0, 0, 0, 0, 0}, delta1[6] = {0, 0, 0, 0, 0, 0};
  int q = 0, ca = 0, cb = 0;
  for (int k = 0; k < count; ++k) {
    const int* r = rows + 8 * (size_t)(first + k);
    q = r[0];
    ca = r[1];
    cb = r[2];
    const int axis = r[3], side = r[4];
    const PatchDesc cq = patches[q];
    const double dtd = axis == 0 ? dt / cq.dx : dt / cq.dy;
    auto add = [&](const double* fl, double* sum) {
      const double* fc = fl + 6 * (size_t)r[5];
      const double* f0 = fl + 6 * (size_t)r[6];
      const double* f1 = fl + 6 * (size_t)r[7];
      for (int f = 0; f < 6; ++f) {
        const double mean = 0.5 * (f0[f] + f1[f]);
        const double d = side == 0 ? dtd * (fc[f] - mean) : dtd * (mean - fc[f]);
        sum[f] = sum[f] + d;
      }
    };
    add(rec, delta);
    add(rec1, delta1);
  }
  const PatchDesc cq = patches[q];
  const size_t PL = (size_t)(cq.nx + 4) * (cq.ny + 4);
  const size_t c = cq.off + (size_t)ca * (cq.ny + 4) + cb;
  for (int f = 0; f < 6; ++f) {
    const double reg = 0.5 * delta1[f];
    dst[c + f * PL] = dst[c + f * PL] + (reg + 0.5 * delta[f]);
  }
}

inline int launch_reflux2d_both(double* dst, const void* patches, const int32_t* rows,
                                const int32_t* groups, int ngroups, const double* rec1,
                                const double* rec2, const double* dt, hipStream_t s) {
  if (ngroups < 0) return WB_EINVAL;
  if (ngroups == 0) return 0;
  if (rows == nullptr || groups == nullptr || rec1 == nullptr || rec2 == nullptr)
    return WB_EINVAL;
  hipLaunchKernelGGL(k_reflux2d_recorded, dim3((unsigned)((ngroups + 63) / 64)), dim3(64), 0, s,
                     dst, (const PatchDesc*)patches, (const int*)rows, (const int*)groups,
                     ngroups, rec1, rec2, dt);
  return wb_last_error();
}

}  // namespace wb
