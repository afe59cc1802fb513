// The fused 2D stage kernel: one launch runs a whole forward-Euler stage of
// MHD2D (primitive variables, MUSCL + HLLD along x and y, the CT-contact nodal
// E_z, the wall faces, the update of the 6 cell fields and of A_z, with the Heun
// average fused into stage 2).
//
// A workgroup owns TY cells along y (the contiguous axis) and marches along x
// over NXS cells.  Thread t maps to the face column ly = t - 1 in [-1, TY]; it
// keeps the x-sweep's four planes of its column in registers, the current
// primitive row sits in LDS for the y-sweep, and the y-face fluxes and the cell
// E_z of the last two rows sit in LDS rings for the CT-contact average.
// P is the arithmetic of the reconstruction and the HLLD flux; the rest is FP64.
// NI adds the resistive and heat-conduction face terms (FP64): T, B_x, B_y, B_z of
// the last three rows and J_z at the nodes of the last two sit in LDS rings.
#pragma once

#include "mhd_device.hpp"
#include "mhd_fast.hpp"
#include "nonideal.hpp"
#include "wbmhd_common.hpp"

namespace wb {

template <int TY>
struct Row {
  static constexpr int FY = TY + 2;  // face columns
  static constexpr int PY = TY + 4;  // primitive columns
  static constexpr int THREADS = ((FY + 63) / 64) * 64;
};

template <typename P, int TY, bool NI = false>
struct Stage2Shared {
  using T = Row<TY>;
  P prim[2][8][T::PY];
  double ecc[3][T::FY];  // cell E_z of rows f-1, f, f+1
  P fx[2][T::FY];     // x-face f: rows 0 (mass) and 5 (B_y)
  P fy[2][7][T::FY];  // y-faces of rows f-1 and f
  // NI: T = p / rho, B_x, B_y, B_z of rows (row % 3); J_z of node rows (row % 2) at
  // node columns jy0 - 1 + c, c in [0, FY]
  double cell[NI ? 3 : 1][4][NI ? T::PY : 1];
  double jz[NI ? 2 : 1][NI ? T::FY + 1 : 1];
};

struct Stage2Args {
  const double* in;
  double* out;
  const double* dt;
  int* bad;
  int nx, ny, X, Y;
  double dx, dy, b0x, b0y, gamma;
  int limiter, heun, nxs;
  int walls;  // faces that are walls: 1 x-low, 2 x-high, 4 y-low, 8 y-high
  Diffusion diff;  // used by the NI kernels only
};

// LIM: the compile-time slope limiter of the branch-free FP32 flux (mixed path),
// or -1 for the parity path's runtime limiter.  EARLY: which stored values the CT and
// the update start from are loaded before the row's second barrier rather than where
// they are used (bit 0: those of `in`, bit 1: those of `out` for the Heun average), so
// that their latency overlaps the barrier, at the price of the registers holding them.
template <typename P, int TY, int LIM, bool NI = false, int EARLY = 3>
__device__ __forceinline__ void stage2d_body(const Stage2Args& a, int tile, int seg,
                                             Stage2Shared<P, TY, NI>& sh) {
  using T = Row<TY>;
  const int tid = threadIdx.x;
  const int X = a.X, Y = a.Y, YN = Y + 1;
  const int jy0 = 2 + tile * TY;
  const int ix0 = 2 + seg * a.nxs;
  const int tyv = min(TY, a.ny + 2 - jy0);
  const int nxv = min(a.nxs, a.nx + 2 - ix0);
  const bool last_y = jy0 + TY >= a.ny + 2, last_x = ix0 + a.nxs >= a.nx + 2;
  const bool fthread = tid < T::FY;
  const int ly = tid - 1, gy = jy0 + ly;
  const size_t PL = (size_t)X * Y;
  const double* az = a.in + 6 * PL;
  const double dt = *a.dt;
  const double dtdx = dt / a.dx, dtdy = dt / a.dy;
  const bool xlo = a.walls & 1, xhi = a.walls & 2, ylo = a.walls & 4, yhi = a.walls & 8;

  P q[7][4];
  P fprev[6], fcur[6];
  double dyy[6];
#pragma unroll
  for (int c = 0; c < 7; ++c)
#pragma unroll
    for (int s = 0; s < 4; ++s) q[c][s] = P(0);
#pragma unroll
  for (int r = 0; r < 6; ++r) { fprev[r] = fcur[r] = P(0); dyy[r] = 0.0; }

  auto AZ = [&](int i, int j) -> double { return az[(size_t)i * YN + j]; };
  // mhd_primitive_pyccel.py:51-86 for row `pl`
  auto primitive_row = [&](int pl, int buf) {
    const int i = min(max(pl, 0), X - 1);
    for (int idx = tid; idx < T::PY; idx += T::THREADS) {
      const int j = min(max(jy0 + idx - 2, 0), Y - 1);
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
      sh.prim[buf][0][idx] = P(r);
      sh.prim[buf][1][idx] = P(u);
      sh.prim[buf][2][idx] = P(v);
      sh.prim[buf][3][idx] = P(w);
      sh.prim[buf][4][idx] = P(p);
      sh.prim[buf][5][idx] = P(bx);
      sh.prim[buf][6][idx] = P(by);
      sh.prim[buf][7][idx] = P(b);
      if (idx >= 1 && idx <= T::FY) sh.ecc[pl % 3][idx - 1] = bx * v - by * u;
      if constexpr (NI) {
        double* cr = &sh.cell[pl % 3][0][0];
        cr[idx] = p / r;
        cr[T::PY + idx] = bx;
        cr[2 * T::PY + idx] = by;
        cr[3 * T::PY + idx] = b;
        if (idx <= T::FY) {  // J_z at node (pl, jy0 - 1 + idx), inside the haloed block
          const int I = min(max(pl, 1), X - 1), J = min(max(jy0 - 1 + idx, 1), Y - 1);
          sh.jz[pl & 1][idx] = node_current_az(az, YN, I, J, a.dx, a.dy, a.b0x, a.b0y);
        }
      }
      if (idx >= 2 && idx < 2 + tyv && pl >= ix0 && pl < ix0 + nxv) {
        const bool is_bad = !(r > 0.0 && p > 0.0) ||
                            !(finite(r) && finite(u) && finite(v) && finite(w) && finite(p) &&
                              finite(bx) && finite(by) && finite(b));
        if (is_bad) atomicAdd(a.bad, 1);
      }
    }
  };
  auto push_column = [&](int buf) {
    if (!fthread) return;
    constexpr int comp[7] = {0, 1, 2, 3, 4, 6, 7};
#pragma unroll
    for (int c = 0; c < 7; ++c) {
      q[c][0] = q[c][1];
      q[c][1] = q[c][2];
      q[c][2] = q[c][3];
      q[c][3] = sh.prim[buf][comp[c]][tid + 1];
    }
  };

  primitive_row(ix0 - 2, (ix0 - 2) & 1);
  __syncthreads();
  push_column((ix0 - 2) & 1);
  __syncthreads();
  primitive_row(ix0 - 1, (ix0 - 1) & 1);
  __syncthreads();
  push_column((ix0 - 1) & 1);

  constexpr int XR[6] = {0, 1, 2, 3, 4, 6};  // x-sweep rows of the 6 fields (mhd.py:134)
  constexpr int YR[6] = {0, 2, 1, 3, 4, 6};
  for (int f = ix0 - 1; f <= ix0 + nxv; ++f) {
    __syncthreads();
    primitive_row(f + 1, (f + 1) & 1);
    __syncthreads();
    push_column((f + 1) & 1);
    const int cur = f & 1;
    if (fthread) {
      const int jj = min(max(gy, 0), Y - 1);
      if (f >= ix0) {  // x-face f of the column
        P fl[7];
        const double bnd = (AZ(f, jj + 1) - AZ(f, jj)) / a.dy + a.b0x;
        const P bn = P(bnd);
        face_solve<P, LIM>(q, a.limiter, bn, P(a.gamma), P(1.0 / (a.gamma - 1.0)), fl);
        if constexpr (NI) {  // mhd.py:868-871, cells L = f - 1 and R = f of column gy
          const double(*cl)[T::PY] = sh.cell[(f + 2) % 3];
          const double(*cr)[T::PY] = sh.cell[f % 3];
          const int c = tid + 1;
          double e4 = double(fl[4]), b6 = double(fl[6]);
          if (a.diff.conduction())
            conductive_face(bnd, cl[2][c], cr[2][c], cl[3][c], cr[3][c], cl[0][c], cr[0][c],
                            limited_gradient(cl[0][c - 1], cl[0][c], cl[0][c + 1], a.dy),
                            limited_gradient(cr[0][c - 1], cr[0][c], cr[0][c + 1], a.dy), a.dx,
                            a.diff, e4);
          if (a.diff.eta != 0.0)
            resistive_xface(cl[3][c], cr[3][c], cl[2][c], cr[2][c], sh.jz[f & 1][tid],
                            sh.jz[f & 1][tid + 1], a.dx, a.diff.eta, e4, b6);
          fl[4] = P(e4);
          fl[6] = P(b6);
        }
        // walls: nothing but the pressure crosses (mhd.py:873-886), after the CT
        const bool wall = (xlo && f == 2) || (xhi && f == a.nx + 2);
#pragma unroll
        for (int r = 0; r < 6; ++r) {
          fprev[r] = fcur[r];
          const int row = XR[r];
          fcur[r] = (wall && row != 1 && row != 5) ? P(0) : fl[row];
        }
        sh.fx[0][tid] = fl[0];
        sh.fx[1][tid] = fl[5];
      }
      if (tid >= 1) {  // y-face below the column's cell: cells ly-2 .. ly+1
        constexpr int comp[7] = {0, 2, 1, 3, 4, 5, 7};
        P qq[7][4];
#pragma unroll
        for (int c = 0; c < 7; ++c)
#pragma unroll
          for (int s = 0; s < 4; ++s) qq[c][s] = sh.prim[cur][comp[c]][tid + 1 + s - 2];
        P fl[7];
        const int ii = min(max(f, 0), X - 1);
        const double bnd = -(AZ(ii + 1, jj) - AZ(ii, jj)) / a.dx + a.b0y;
        const P bn = P(bnd);
        face_solve<P, LIM>(qq, a.limiter, bn, P(a.gamma), P(1.0 / (a.gamma - 1.0)), fl);
        if constexpr (NI) {  // cells L = gy - 1 and R = gy of row f
          const double(*cm)[T::PY] = sh.cell[(f + 2) % 3];
          const double(*c0)[T::PY] = sh.cell[f % 3];
          const double(*cp)[T::PY] = sh.cell[(f + 1) % 3];
          const int c = tid + 1;
          double e4 = double(fl[4]), b6 = double(fl[6]);
          if (a.diff.conduction())
            conductive_face(bnd, c0[1][c - 1], c0[1][c], c0[3][c - 1], c0[3][c], c0[0][c - 1],
                            c0[0][c], limited_gradient(cm[0][c - 1], c0[0][c - 1], cp[0][c - 1], a.dx),
                            limited_gradient(cm[0][c], c0[0][c], cp[0][c], a.dx), a.dy, a.diff, e4);
          if (a.diff.eta != 0.0)
            resistive_yface(c0[3][c - 1], c0[3][c], c0[1][c - 1], c0[1][c], sh.jz[f & 1][tid],
                            sh.jz[(f + 1) & 1][tid], a.dy, a.diff.eta, e4, b6);
          fl[4] = P(e4);
          fl[6] = P(b6);
        }
#pragma unroll
        for (int r = 0; r < 7; ++r) sh.fy[cur][r][tid] = fl[r];
      }
    }
    // the stored values of A_z at node (f, gy) and of the cell (f - 1, gy) that the CT
    // and the update start from, and their values in `out` for the Heun average
    const bool in_y = ly >= 0 && (ly < tyv || (ly == tyv && last_y));
    const bool cell_y = ly >= 0 && ly < tyv;
    const bool node_x = f >= ix0 && (f < ix0 + nxv || (f == ix0 + nxv && last_x));
    const bool do_node = fthread && node_x && in_y;
    const bool do_cell = fthread && cell_y && f - 1 >= ix0 && f - 1 < ix0 + nxv;
    const size_t onode = 6 * PL + (size_t)f * YN + gy;
    const size_t ocell = (size_t)(f - 1) * Y + gy;
    double vin[7], vout[7];  // A_z, then the 6 cell fields
#pragma unroll
    for (int r = 0; r < 7; ++r) vin[r] = vout[r] = 0.0;
    if constexpr (EARLY & 1) {
      if (do_node) vin[0] = a.in[onode];
      if (do_cell) {
#pragma unroll
        for (int r = 0; r < 6; ++r) vin[1 + r] = a.in[r * PL + ocell];
      }
    }
    if ((EARLY & 2) && a.heun) {
      if (do_node) vout[0] = a.out[onode];
      if (do_cell) {
#pragma unroll
        for (int r = 0; r < 6; ++r) vout[1 + r] = a.out[r * PL + ocell];
      }
    }
    __syncthreads();
    if (fthread) {
      const int prv = cur ^ 1;
      // E_z at (x-node f, y-node ly) (ct_contact_pyccel.py:35-61), walls zero it
      if (do_node) {
        double e = ct_node(double(sh.fx[1][tid - 1]), double(sh.fx[1][tid]),
                           double(sh.fy[prv][5][tid]), double(sh.fy[cur][5][tid]),
                           double(sh.fx[0][tid - 1]), double(sh.fx[0][tid]),
                           double(sh.fy[prv][0][tid]), double(sh.fy[cur][0][tid]),
                           sh.ecc[(f + 2) % 3][tid - 1], sh.ecc[f % 3][tid - 1],
                           sh.ecc[(f + 2) % 3][tid], sh.ecc[f % 3][tid]);
        if constexpr (NI) {
          if (a.diff.eta != 0.0) e = e + a.diff.eta * sh.jz[f & 1][tid];
        }
        if ((xlo && f == 2) || (xhi && f == a.nx + 2) || (ylo && gy == 2) ||
            (yhi && gy == a.ny + 2))
          e = 0.0;
        double v = ((EARLY & 1) ? vin[0] : a.in[onode]) - dt * e;
        if (a.heun) v = (v + ((EARLY & 2) ? vout[0] : a.out[onode])) * 0.5;
        a.out[onode] = v;
      }
      if (cell_y) {
        if (do_cell) {
#pragma unroll
          for (int r = 0; r < 6; ++r) {
            const double dfx = double(fcur[r]) - double(fprev[r]);
            const size_t o = r * PL + ocell;
            double v = ((EARLY & 1) ? vin[1 + r] : a.in[o]) - (dtdx * dfx + dyy[r]);
            if (a.heun) v = (v + ((EARLY & 2) ? vout[1 + r] : a.out[o])) * 0.5;
            a.out[o] = v;
          }
        }
        if (f >= ix0 && f < ix0 + nxv) {
          const bool wlo = ylo && gy == 2, whi = yhi && gy + 1 == a.ny + 2;
#pragma unroll
          for (int r = 0; r < 6; ++r) {
            const int row = YR[r];
            const bool zeroable = row != 1 && row != 5;
            const double hi = (whi && zeroable) ? 0.0 : double(sh.fy[cur][row][tid + 1]);
            const double lo = (wlo && zeroable) ? 0.0 : double(sh.fy[cur][row][tid]);
            dyy[r] = dtdy * (hi - lo);
          }
        }
      }
    }
  }
}

template <typename P, int TY, int LIM, bool NI = false>
__global__ void __launch_bounds__((Row<TY>::THREADS)) __attribute__((amdgpu_waves_per_eu(1, 3)))
k_stage2d(Stage2Args a) {
  __shared__ Stage2Shared<P, TY, NI> sh;
  stage2d_body<P, TY, LIM, NI>(a, blockIdx.x, blockIdx.y, sh);
}


// All tiles of all patches in one persistent launch: each workgroup takes the next
// work item from a counter until the list is exhausted (the counter is zeroed by
// k_zero_i32 before the launch).
template <typename P, int TY, int LIM, bool NI = false>
__global__ void __launch_bounds__((Row<TY>::THREADS)) __attribute__((amdgpu_waves_per_eu(1, 3)))
k_stage2d_batched(Stage2Args base, const PatchDesc* __restrict__ patches,
                  const WorkItem* __restrict__ items, int nitems, int* __restrict__ counter,
                  double* __restrict__ out_base) {
  __shared__ Stage2Shared<P, TY, NI> sh;
  __shared__ int next;
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(counter, 1);
    __syncthreads();
    const int it = next;
    __syncthreads();
    if (it >= nitems) break;
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
    stage2d_body<P, TY, LIM, NI>(a, w.tile, w.seg, sh);
    __syncthreads();
  }
}

// The flux of one face of a patch from its (ghost-filled) stage input, computed as
// the stage kernels compute it: primitive variables in FP64, reconstruction and
// HLLD in P.  axis 0: x-face `fa` (low face of cell fa) at row fb; axis 1: y-face.
template <typename P, int LIM>
__device__ void face_flux2d(const double* in, int X, int Y, double dx, double dy, double b0x,
                            double b0y, double gamma, int limiter, const Diffusion& diff, int axis,
                            int fa, int fb, double* out7) {
  const size_t PL = (size_t)X * Y;
  const double* az = in + 6 * PL;
  const int YN = Y + 1;
  auto AZ = [&](int i, int j) -> double { return az[(size_t)i * YN + j]; };
  P q[7][4];
  for (int s = 0; s < 4; ++s) {
    const int i = axis == 0 ? fa - 2 + s : fa;
    const int j = axis == 0 ? fb : fb - 2 + s;
    const size_t c = (size_t)i * Y + j;
    const double bx0 = (AZ(i, j + 1) - AZ(i, j)) / dy + b0x;
    const double bx1 = (AZ(i + 1, j + 1) - AZ(i + 1, j)) / dy + b0x;
    const double by0 = -(AZ(i + 1, j) - AZ(i, j)) / dx + b0y;
    const double by1 = -(AZ(i + 1, j + 1) - AZ(i, j + 1)) / dx + b0y;
    const double r = in[c];
    const double u = in[PL + c] / r, v = in[2 * PL + c] / r, w = in[3 * PL + c] / r;
    const double bx = 0.5 * (bx0 + bx1);
    const double by = 0.5 * (by0 + by1);
    const double b = in[5 * PL + c];
    const double p = (gamma - 1.0) * (in[4 * PL + c] - 0.5 * r * (u * u + v * v + w * w) -
                                      0.5 * (bx * bx + by * by + b * b));
    const double pr[8] = {r, u, v, w, p, bx, by, b};
    const int comp0[7] = {0, 1, 2, 3, 4, 6, 7};
    const int comp1[7] = {0, 2, 1, 3, 4, 5, 7};
    for (int c7 = 0; c7 < 7; ++c7) q[c7][s] = P(pr[axis == 0 ? comp0[c7] : comp1[c7]]);
  }
  const double bn = axis == 0 ? (AZ(fa, fb + 1) - AZ(fa, fb)) / dy + b0x
                              : -(AZ(fa + 1, fb) - AZ(fa, fb)) / dx + b0y;
  P f[7];
  face_solve<P, LIM>(q, limiter, P(bn), P(gamma), P(1.0 / (gamma - 1.0)), f);
  for (int c = 0; c < 7; ++c) out7[c] = double(f[c]);
  if (!diff.conduction() && diff.eta == 0.0) return;
  // the non-ideal terms (mhd.py:868-871) from T, B of the cells around the face
  auto cell = [&](int i, int j, double out4[4]) {
    const size_t c = (size_t)i * Y + j;
    const double bx0 = (AZ(i, j + 1) - AZ(i, j)) / dy + b0x;
    const double bx1 = (AZ(i + 1, j + 1) - AZ(i + 1, j)) / dy + b0x;
    const double by0 = -(AZ(i + 1, j) - AZ(i, j)) / dx + b0y;
    const double by1 = -(AZ(i + 1, j + 1) - AZ(i, j + 1)) / dx + b0y;
    const double r = in[c];
    const double u = in[PL + c] / r, v = in[2 * PL + c] / r, w = in[3 * PL + c] / r;
    const double bx = 0.5 * (bx0 + bx1);
    const double by = 0.5 * (by0 + by1);
    const double b = in[5 * PL + c];
    const double p = (gamma - 1.0) * (in[4 * PL + c] - 0.5 * r * (u * u + v * v + w * w) -
                                      0.5 * (bx * bx + by * by + b * b));
    out4[0] = p / r;
    out4[1] = bx;
    out4[2] = by;
    out4[3] = b;
  };
  double e4 = out7[4], b6 = out7[6];
  // L and R beside the face, and the cells before (m) and after (p) each along it
  const int di = axis == 0 ? 1 : 0, dj = 1 - di;
  double L[4], R[4], Lm[4], Lp[4], Rm[4], Rp[4];
  cell(fa - di, fb - dj, L);
  cell(fa, fb, R);
  if (diff.conduction()) {
    cell(fa - di - dj, fb - dj - di, Lm);
    cell(fa - di + dj, fb - dj + di, Lp);
    cell(fa - dj, fb - di, Rm);
    cell(fa + dj, fb + di, Rp);
    const double ha = axis == 0 ? dy : dx;  // the spacing along the face
    conductive_face(bn, axis == 0 ? L[2] : L[1], axis == 0 ? R[2] : R[1], L[3], R[3], L[0], R[0],
                    limited_gradient(Lm[0], L[0], Lp[0], ha), limited_gradient(Rm[0], R[0], Rp[0], ha),
                    axis == 0 ? dx : dy, diff, e4);
  }
  if (diff.eta != 0.0) {
    const double j0 = node_current_az(az, YN, fa, fb, dx, dy, b0x, b0y);
    const double j1 = node_current_az(az, YN, fa + dj, fb + di, dx, dy, b0x, b0y);
    if (axis == 0)
      resistive_xface(L[3], R[3], L[2], R[2], j0, j1, dx, diff.eta, e4, b6);
    else
      resistive_yface(L[3], R[3], L[1], R[1], j0, j1, dy, diff.eta, e4, b6);
  }
  out7[4] = double(P(e4));
  out7[6] = double(P(b6));
}

// The flux register and refluxing (patchwork FluxRegister.update/reflux as MHD2D.step
// calls them, mhd.py:911-919,989-990): per uncovered coarse cell next to finer patches,
// d = sum over its coarse-fine faces (x-high, x-low, y-high, y-low) of dt/h (F_c - mean)
// (high face) or dt/h (mean - F_c) (low face), mean = 0.5 (F_f0 + F_f1), the fluxes
// recomputed from the stage input.  record (stage 1): reg = 0.5 d; otherwise (stage 2,
// after the Heun average): cell += reg + 0.5 d, the coarse fluxes of both stages traded
// for the fine ones at the 0.5 dt weight each stage recorded them with.  rows (13 int64
// each): q, ca, cb, axis, side, face_a, face_b, p0, fa0, fb0, p1, fa1, fb1; groups[g] =
// (first row, row count) of coarse cell g; reg holds 6 doubles per group.
// Deterministic: one thread per cell.
template <typename P, int LIM>
__global__ void __launch_bounds__(64) k_reflux2d(const double* __restrict__ src, double* __restrict__ dst,
                           const PatchDesc* __restrict__ patches, const long long* __restrict__ rows,
                           const int* __restrict__ groups, int ngroups, double b0x, double b0y,
                           double gamma, int limiter, Diffusion diff, const double* __restrict__ dtp,
                           int record, double* __restrict__ reg) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  constexpr int XR[6] = {0, 1, 2, 3, 4, 6};
  constexpr int YR[6] = {0, 2, 1, 3, 4, 6};
  const double dt = *dtp;
  const int first = groups[2 * g], count = groups[2 * g + 1];
  double delta[6] = {0, 0, 0, 0, 0, 0};
  long long q = 0, ca = 0, cb = 0;
  for (int k = 0; k < count; ++k) {
    const long long* r = rows + 13 * (size_t)(first + k);
    q = r[0];
    ca = r[1];
    cb = r[2];
    const int axis = (int)r[3], side = (int)r[4];
    const PatchDesc cq = patches[q];
    const double dtd = axis == 0 ? dt / cq.dx : dt / cq.dy;
    double fc[7], f0[7], f1[7];
    face_flux2d<P, LIM>(src + cq.off, cq.nx + 4, cq.ny + 4, cq.dx, cq.dy, b0x, b0y, gamma, limiter,
                        diff, axis, (int)r[5], (int)r[6], fc);
    const PatchDesc p0 = patches[r[7]], p1 = patches[r[10]];
    face_flux2d<P, LIM>(src + p0.off, p0.nx + 4, p0.ny + 4, p0.dx, p0.dy, b0x, b0y, gamma, limiter,
                        diff, axis, (int)r[8], (int)r[9], f0);
    face_flux2d<P, LIM>(src + p1.off, p1.nx + 4, p1.ny + 4, p1.dx, p1.dy, b0x, b0y, gamma, limiter,
                        diff, axis, (int)r[11], (int)r[12], f1);
    for (int f = 0; f < 6; ++f) {
      const int row = axis == 0 ? XR[f] : YR[f];
      const double mean = 0.5 * (f0[row] + f1[row]);
      const double d = side == 0 ? dtd * (fc[row] - mean) : dtd * (mean - fc[row]);
      delta[f] = delta[f] + d;
    }
  }
  double* kept = reg + 6 * (size_t)g;
  if (record) {
    for (int f = 0; f < 6; ++f) kept[f] = 0.5 * delta[f];
    return;
  }
  const PatchDesc cq = patches[q];
  const size_t PL = (size_t)(cq.nx + 4) * (cq.ny + 4);
  const size_t c = cq.off + (size_t)ca * (cq.ny + 4) + cb;
  for (int f = 0; f < 6; ++f) dst[c + f * PL] = dst[c + f * PL] + (kept[f] + 0.5 * delta[f]);
}

template <typename P, int TY, int LIM = -1, bool NI = false>
inline int launch_stage2d(const Stage2Args& a, hipStream_t s) {
  const int tiles = (a.ny + TY - 1) / TY;
  const int segs = (a.nx + a.nxs - 1) / a.nxs;
  hipLaunchKernelGGL((k_stage2d<P, TY, LIM, NI>), dim3(tiles, segs), dim3(Row<TY>::THREADS), 0, s,
                     a);
  return wb_last_error();
}

}  // namespace wb
