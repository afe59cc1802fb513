// The fused 3D stage kernel: one launch runs a whole forward-Euler stage of the
// ideal-MHD step (primitive variables, MUSCL + HLLD along x, y and z, the
// CT-contact edge field, the conservative update of the 5 cell fields and the
// update of the edge potential, with the Heun average fused into stage 2).
//
// Work decomposition (2.5D): a workgroup owns a TY x TZ tile of cells in the
// (y, z) plane and marches along x over a segment of NXS cells.  Its threads map
// to the (TY+2) x (TZ+2) "face columns" (local y, z in [-1, T]).  Each thread keeps
// the x-sweep's four planes of its column in registers (a ring shifted once per
// plane), so the x-faces need no LDS; the current primitive plane (T+4)^2 is
// staged in LDS for the y- and z-sweeps; the y/z-face fluxes and the cell
// electric field of the last two planes sit in LDS rings for the CT-contact
// average; the x-face fluxes of the column stay in registers for the update.
// Every face flux a tile needs is computed once per tile; faces on the tile
// border are computed by both neighbouring tiles.
//
// P is the arithmetic of the reconstruction and the HLLD flux (double: the
// parity path; float: the mixed path).  The primitive conversion (E -> p), the
// cell electric field, the CT average, the face field from the potential and
// the update are FP64 in both.
#pragma once

#include "mhd_device.hpp"
#include "wbmhd_common.hpp"

namespace wb {

struct Grid3 {
  int X, Y, Z;     // haloed extents
  int nx, ny, nz;  // interior cells
};

__device__ __forceinline__ size_t cidx(const Grid3& g, int i, int j, int k) {
  return ((size_t)i * g.Y + j) * g.Z + k;
}

template <int TY, int TZ>
struct Tile {
  static constexpr int FY = TY + 2, FZ = TZ + 2;  // face columns
  static constexpr int PY = TY + 4, PZ = TZ + 4;  // primitive columns
  static constexpr int NF = FY * FZ;
  static constexpr int NP = PY * PZ;
  static constexpr int THREADS = ((NF + 63) / 64) * 64;
};

template <typename P, int TY, int TZ>
struct StageShared {
  using T = Tile<TY, TZ>;
  P prim[2][8][T::NP];  // primitive planes f and f+1
  double ex[2][T::NF];  // cell E_x of planes f, f+1
  double ey[3][T::NF];  // cell E_y of planes f-1, f, f+1
  double ez[3][T::NF];
  P fx[3][T::NF];       // x-face f: rows 0 (mass), 5 (B_y), 6 (B_z)
  P fy[2][7][T::NF];    // y-faces of planes f-1, f (face below the column's cell)
  P fz[2][7][T::NF];
};

struct StageArgs {
  const double* in;
  double* out;
  const double* dt;
  int* bad;
  Grid3 g;
  double dx, dy, dz, b0x, b0y, b0z, gamma;
  int limiter, heun, nxs;
  size_t off_ax, off_ay, off_az;
};

// Face-normal fields from the edge potential (mhd3d.py:361-378), the same
// expressions as the split kernels so that both paths agree bit for bit.
struct EdgePot {
  const double *ax, *ay, *az;
  int Y, Z;
  __device__ double AX(int i, int j, int k) const { return ax[((size_t)i * (Y + 1) + j) * (Z + 1) + k]; }
  __device__ double AY(int i, int j, int k) const { return ay[((size_t)i * Y + j) * (Z + 1) + k]; }
  __device__ double AZ(int i, int j, int k) const { return az[((size_t)i * (Y + 1) + j) * Z + k]; }
  __device__ double bx(int i, int j, int k, double dy, double dz, double b0) const {
    return (AZ(i, j + 1, k) - AZ(i, j, k)) / dy - (AY(i, j, k + 1) - AY(i, j, k)) / dz + b0;
  }
  __device__ double by(int i, int j, int k, double dz, double dx, double b0) const {
    return (AX(i, j, k + 1) - AX(i, j, k)) / dz - (AZ(i + 1, j, k) - AZ(i, j, k)) / dx + b0;
  }
  __device__ double bz(int i, int j, int k, double dx, double dy, double b0) const {
    return (AY(i + 1, j, k) - AY(i, j, k)) / dx - (AX(i, j + 1, k) - AX(i, j, k)) / dy + b0;
  }
};

template <typename P, int TY, int TZ>
__global__ void __launch_bounds__((Tile<TY, TZ>::THREADS))
k_stage3d(StageArgs a) {
  using T = Tile<TY, TZ>;
  __shared__ StageShared<P, TY, TZ> sh;
  const Grid3 g = a.g;
  const int tid = threadIdx.x;
  const int tiles_z = (g.nz + TZ - 1) / TZ;
  const int by = blockIdx.x / tiles_z, bz = blockIdx.x % tiles_z;
  const int jy0 = 2 + by * TY, kz0 = 2 + bz * TZ;       // first interior cell of the tile
  const int ix0 = 2 + blockIdx.y * a.nxs;               // first plane of the segment
  const int tyv = min(TY, g.ny + 2 - jy0), tzv = min(TZ, g.nz + 2 - kz0);
  const int nxv = min(a.nxs, g.nx + 2 - ix0);
  const bool last_y = jy0 + TY >= g.ny + 2, last_z = kz0 + TZ >= g.nz + 2;
  const bool last_x = ix0 + a.nxs >= g.nx + 2;
  // this thread's face column (local y, z in [-1, T])
  const bool fthread = tid < T::NF;
  const int fa = tid / T::FZ, fb = tid % T::FZ;
  const int ly = fa - 1, lz = fb - 1;
  const int gy = jy0 + ly, gz = kz0 + lz;
  const EdgePot A{a.in + a.off_ax, a.in + a.off_ay, a.in + a.off_az, g.Y, g.Z};
  const size_t PL = (size_t)g.X * g.Y * g.Z;
  const double dt = *a.dt;
  const double dtdx = dt / a.dx, dtdy = dt / a.dy, dtdz = dt / a.dz;

  // register state of the column
  P q[7][4];                       // x-frame primitives of planes f-2 .. f+1
  P fprev[5], fcur[5];             // x-face fluxes (rows 0-4) of faces f-1 and f
  double dyz[5];                   // dt/dy dFy + ... partial terms of the pending cell
  double dyy[5], dzz[5];
#pragma unroll
  for (int c = 0; c < 7; ++c)
#pragma unroll
    for (int s = 0; s < 4; ++s) q[c][s] = P(0);
#pragma unroll
  for (int r = 0; r < 5; ++r) { fprev[r] = fcur[r] = P(0); dyy[r] = dzz[r] = dyz[r] = 0.0; }

  // -- primitive plane `pl` into LDS buffer `buf`, cell E into the rings ----------
  auto primitive_plane = [&](int pl, int buf, bool to_lds) {
    const int pc = min(max(pl, 0), g.X - 1);  // clamp keeps every load in bounds
    for (int idx = tid; idx < T::NP; idx += T::THREADS) {
      const int py = idx / T::PZ, pz = idx % T::PZ;
      const int j = min(max(jy0 + py - 2, 0), g.Y - 1), k = min(max(kz0 + pz - 2, 0), g.Z - 1);
      const size_t c = cidx(g, pc, j, k);
      const double bx = 0.5 * (A.bx(pc, j, k, a.dy, a.dz, a.b0x) + A.bx(pc + 1, j, k, a.dy, a.dz, a.b0x));
      const double byv = 0.5 * (A.by(pc, j, k, a.dz, a.dx, a.b0y) + A.by(pc, j + 1, k, a.dz, a.dx, a.b0y));
      const double bzv = 0.5 * (A.bz(pc, j, k, a.dx, a.dy, a.b0z) + A.bz(pc, j, k + 1, a.dx, a.dy, a.b0z));
      const double r = a.in[c];
      const double u = a.in[PL + c] / r, v = a.in[2 * PL + c] / r, w = a.in[3 * PL + c] / r;
      const double p = (a.gamma - 1.0) * (a.in[4 * PL + c] - 0.5 * r * (u * u + v * v + w * w) -
                                          0.5 * (bx * bx + byv * byv + bzv * bzv));
      if (to_lds) {
        sh.prim[buf][0][idx] = P(r);
        sh.prim[buf][1][idx] = P(u);
        sh.prim[buf][2][idx] = P(v);
        sh.prim[buf][3][idx] = P(w);
        sh.prim[buf][4][idx] = P(p);
        sh.prim[buf][5][idx] = P(bx);
        sh.prim[buf][6][idx] = P(byv);
        sh.prim[buf][7][idx] = P(bzv);
      }
      // the face-column region [-1, T] holds the cell electric field
      const int fy_ = py - 1, fz_ = pz - 1;
      if (fy_ >= 0 && fy_ < T::FY && fz_ >= 0 && fz_ < T::FZ) {
        const int f = fy_ * T::FZ + fz_;
        sh.ex[pl & 1][f] = byv * w - bzv * v;
        sh.ey[(pl + 3) % 3][f] = bzv * u - bx * w;
        sh.ez[(pl + 3) % 3][f] = bx * v - byv * u;
      }
      // bad cells: only the tile's own interior cells of the segment's planes
      if (py >= 2 && py < 2 + tyv && pz >= 2 && pz < 2 + tzv && pl >= ix0 && pl < ix0 + nxv) {
        const bool ok = finite(r) && finite(u) && finite(v) && finite(w) && finite(p) &&
                        finite(bx) && finite(byv) && finite(bzv) && r > 0.0 && p > 0.0;
        if (!ok) atomicAdd(a.bad, 1);
      }
    }
  };
  // the x-frame primitives of the own column, pushed into the register ring
  auto push_column = [&](int buf) {
    if (!fthread) return;
    const int idx = (fa + 1) * T::PZ + (fb + 1);
    constexpr int comp[7] = {0, 1, 2, 3, 4, 6, 7};
#pragma unroll
    for (int c = 0; c < 7; ++c) {
      q[c][0] = q[c][1];
      q[c][1] = q[c][2];
      q[c][2] = q[c][3];
      q[c][3] = sh.prim[buf][comp[c]][idx];
    }
  };

  // prologue: planes ix0-2 (ring only) and ix0-1 (ring + LDS)
  {
    const int pl = ix0 - 2;
    primitive_plane(pl, pl & 1, true);
    __syncthreads();
    push_column(pl & 1);
    __syncthreads();
    primitive_plane(pl + 1, (pl + 1) & 1, true);
    __syncthreads();
    push_column((pl + 1) & 1);
  }

  for (int f = ix0 - 1; f <= ix0 + nxv; ++f) {
    __syncthreads();
    // 1. primitive plane f+1
    primitive_plane(f + 1, (f + 1) & 1, true);
    __syncthreads();
    push_column((f + 1) & 1);
    const int cur = f & 1;
    if (fthread) {
      // 3. x-face f of the column (lower face of cell f): planes f-2 .. f+1
      if (f >= ix0) {
        P L[7], R[7], fl[7];
        reconstruct<P>(q, a.limiter, L, R);
        const P bn = P(A.bx(f, min(max(gy, 0), g.Y - 1), min(max(gz, 0), g.Z - 1), a.dy, a.dz, a.b0x));
        hlld<P>(L, R, bn, P(a.gamma), fl);
#pragma unroll
        for (int r = 0; r < 5; ++r) {
          fprev[r] = fcur[r];
          fcur[r] = fl[r];
        }
        sh.fx[0][tid] = fl[0];
        sh.fx[1][tid] = fl[5];
        sh.fx[2][tid] = fl[6];
      }
      // 4. y- and z-faces of plane f below the column's cell
      {
        const int pidx = (fa + 1) * T::PZ + (fb + 1);
        // y-face between cells (ly-1) and ly: cells ly-2 .. ly+1 in frame (y, z, x)
        if (fa >= 1) {
          constexpr int comp[7] = {0, 2, 3, 1, 4, 7, 5};
          P qq[7][4];
#pragma unroll
          for (int c = 0; c < 7; ++c)
#pragma unroll
            for (int s = 0; s < 4; ++s) qq[c][s] = sh.prim[cur][comp[c]][pidx + (s - 2) * T::PZ];
          P L[7], R[7], fl[7];
          reconstruct<P>(qq, a.limiter, L, R);
          const int jj = min(max(gy, 0), g.Y - 1), kk = min(max(gz, 0), g.Z - 1);
          const P bn = P(A.by(min(max(f, 0), g.X - 1), jj, kk, a.dz, a.dx, a.b0y));
          hlld<P>(L, R, bn, P(a.gamma), fl);
#pragma unroll
          for (int r = 0; r < 7; ++r) sh.fy[cur][r][tid] = fl[r];
        }
        if (fb >= 1) {
          constexpr int comp[7] = {0, 3, 1, 2, 4, 5, 6};
          P qq[7][4];
#pragma unroll
          for (int c = 0; c < 7; ++c)
#pragma unroll
            for (int s = 0; s < 4; ++s) qq[c][s] = sh.prim[cur][comp[c]][pidx + (s - 2)];
          P L[7], R[7], fl[7];
          reconstruct<P>(qq, a.limiter, L, R);
          const int jj = min(max(gy, 0), g.Y - 1), kk = min(max(gz, 0), g.Z - 1);
          const P bn = P(A.bz(min(max(f, 0), g.X - 1), jj, kk, a.dx, a.dy, a.b0z));
          hlld<P>(L, R, bn, P(a.gamma), fl);
#pragma unroll
          for (int r = 0; r < 7; ++r) sh.fz[cur][r][tid] = fl[r];
        }
      }
    }
    __syncthreads();
    if (fthread) {
      const int prv = cur ^ 1;
      const int e0 = (f + 2) % 3, e1 = (f + 3) % 3;  // planes f-1 and f of the E rings
      const bool in_y = ly >= 0 && (ly < tyv || (ly == tyv && last_y));
      const bool in_z = lz >= 0 && (lz < tzv || (lz == tzv && last_z));
      const bool cell_y = ly >= 0 && ly < tyv, cell_z = lz >= 0 && lz < tzv;
      const int tym = tid - T::FZ, tzm = tid - 1, tymz = tid - T::FZ - 1;
      // 6a. E_x on the x-cell plane f at (y-node ly, z-node lz): plane (a=y, b=z)
      if (f >= ix0 && f < ix0 + nxv && in_y && in_z && ly <= TY && lz <= TZ) {
        const double e = ct_node(
            double(sh.fy[cur][5][tzm]), double(sh.fy[cur][5][tid]),
            double(sh.fz[cur][6][tym]), double(sh.fz[cur][6][tid]),
            double(sh.fy[cur][0][tzm]), double(sh.fy[cur][0][tid]),
            double(sh.fz[cur][0][tym]), double(sh.fz[cur][0][tid]),
            sh.ex[f & 1][tymz], sh.ex[f & 1][tzm], sh.ex[f & 1][tym], sh.ex[f & 1][tid]);
        const size_t o = a.off_ax + ((size_t)f * (g.Y + 1) + gy) * (g.Z + 1) + gz;
        double v = a.in[o] - dt * e;
        if (a.heun) v = (v + a.out[o]) * 0.5;
        a.out[o] = v;
      }
      const bool node_x = f >= ix0 && (f < ix0 + nxv || (f == ix0 + nxv && last_x));
      // 6b. E_y at (x-node f, y-cell ly, z-node lz): plane (a=z, b=x)
      if (node_x && cell_y && in_z) {
        const double e = ct_node(
            double(sh.fz[prv][5][tid]), double(sh.fz[cur][5][tid]),
            double(sh.fx[2][tzm]), double(sh.fx[2][tid]),
            double(sh.fz[prv][0][tid]), double(sh.fz[cur][0][tid]),
            double(sh.fx[0][tzm]), double(sh.fx[0][tid]),
            sh.ey[e0][tzm], sh.ey[e0][tid], sh.ey[e1][tzm], sh.ey[e1][tid]);
        const size_t o = a.off_ay + ((size_t)f * g.Y + gy) * (g.Z + 1) + gz;
        double v = a.in[o] - dt * e;
        if (a.heun) v = (v + a.out[o]) * 0.5;
        a.out[o] = v;
      }
      // 6c. E_z at (x-node f, y-node ly, z-cell lz): plane (a=x, b=y)
      if (node_x && in_y && cell_z) {
        const double e = ct_node(
            double(sh.fx[1][tym]), double(sh.fx[1][tid]),
            double(sh.fy[prv][6][tid]), double(sh.fy[cur][6][tid]),
            double(sh.fx[0][tym]), double(sh.fx[0][tid]),
            double(sh.fy[prv][0][tid]), double(sh.fy[cur][0][tid]),
            sh.ez[e0][tym], sh.ez[e1][tym], sh.ez[e0][tid], sh.ez[e1][tid]);
        const size_t o = a.off_az + ((size_t)f * (g.Y + 1) + gy) * g.Z + gz;
        double v = a.in[o] - dt * e;
        if (a.heun) v = (v + a.out[o]) * 0.5;
        a.out[o] = v;
      }
      // 7. update of cell (f-1, ly, lz): x-faces f-1 and f, y/z terms of plane f-1
      if (cell_y && cell_z) {
        if (f - 1 >= ix0 && f - 1 < ix0 + nxv) {
          const size_t c = cidx(g, f - 1, gy, gz);
#pragma unroll
          for (int r = 0; r < 5; ++r) {
            const double dfx = double(fcur[r]) - double(fprev[r]);
            double v = a.in[r * PL + c] - (dtdx * dfx + dyy[r] + dzz[r]);
            if (a.heun) v = (v + a.out[r * PL + c]) * 0.5;
            a.out[r * PL + c] = v;
          }
        }
        // y/z flux differences of cell (f, ly, lz) for the next iteration
        if (f >= ix0 && f < ix0 + nxv) {
          const int typ = tid + T::FZ, tzp = tid + 1;
          // rows of the y-sweep (rho, m_y, m_z, m_x, E) and z-sweep (rho, m_z, m_x, m_y, E)
          constexpr int yrow[5] = {0, 3, 1, 2, 4};
          constexpr int zrow[5] = {0, 2, 3, 1, 4};
#pragma unroll
          for (int r = 0; r < 5; ++r) {
            dyy[r] = dtdy * (double(sh.fy[cur][yrow[r]][typ]) - double(sh.fy[cur][yrow[r]][tid]));
            dzz[r] = dtdz * (double(sh.fz[cur][zrow[r]][tzp]) - double(sh.fz[cur][zrow[r]][tid]));
          }
        }
      }
    }
  }
}

template <typename P, int TY, int TZ>
inline int launch_stage3d(const StageArgs& a, hipStream_t s) {
  using T = Tile<TY, TZ>;
  const int tiles = ((a.g.ny + TY - 1) / TY) * ((a.g.nz + TZ - 1) / TZ);
  const int segs = (a.g.nx + a.nxs - 1) / a.nxs;
  hipLaunchKernelGGL((k_stage3d<P, TY, TZ>), dim3(tiles, segs), dim3(T::THREADS), 0, s, a);
  return wb_last_error();
}

}  // namespace wb
