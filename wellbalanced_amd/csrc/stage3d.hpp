// The fused 3D stage kernel: one launch runs a whole forward-Euler stage of the
// ideal-MHD step (primitive variables, MUSCL + HLLD along x, y and z, the
// CT-contact edge field, the conservative update of the 5 cell fields and the
// update of the edge potential, with the Heun average fused into stage 2).
//
// Work decomposition (2.5D): a workgroup owns a TY x TZ tile of cells in the
// (y, z) plane and marches along x over a segment of NXS cells.  Its threads map
// to the (TY+2) x (TZ+2) "face columns" (local y, z in [-1, T]).  Each thread keeps
// the x-sweep's four planes of its column in registers (a ring shifted once per
// plane), so the x-faces need no LDS; the current primitive plane (T+4)x(T+4) is
// staged in LDS for the y- and z-sweeps; the y/z-face fluxes and the cell
// electric field of the last planes sit in LDS rings for the CT-contact average;
// the x-face fluxes of the column stay in registers for the update.  Every face
// flux a tile needs is computed once per tile; faces on the tile border are
// computed by both neighbouring tiles.  Per plane: the y/z faces (from the LDS
// plane), then the column's next plane and its x-face (registers), one barrier,
// then the next LDS plane, the CT and the update.  The mixed path's LDS fits four
// workgroups of 256 threads per CU.
//
// P is the arithmetic of the reconstruction and the HLLD flux (double: the
// parity path; float: the mixed path).  The primitive conversion (E -> p), the
// cell electric field, the CT average, the face field from the potential and
// the update are FP64 in both; the mixed path keeps the cell electric field for
// the CT in FP32, as it keeps the face fluxes the CT averages.
#pragma once

#include "mhd_device.hpp"
#include "mhd_fast.hpp"
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
  P prim[8][T::NP];      // the primitive plane f (the y/z sweeps read it)
  P ex[2][T::NF];        // cell E_x of planes f, f+1
  P ey[3][T::NF];        // cell E_y of planes f-1, f, f+1
  P ez[3][T::NF];
  P fx[3][T::NF];        // x-face f: rows 0 (mass), 5 (B_y), 6 (B_z)
  P fy[5][T::NF];        // rows 1-5 of the y-faces of plane f (face below the column's cell)
  P fz[5][T::NF];        // rows 1-4 and 6 of the z-faces of plane f
  P fyc[2][2][T::NF];    // rows 0 (mass) and 6 (B_x) of the y-faces of planes f-1, f
  P fzc[2][2][T::NF];    // rows 0 (mass) and 5 (B_x) of the z-faces of planes f-1, f
};

struct StageArgs {
  const double* in;
  double* out;
  const double* dt;
  int* bad;
  Grid3 g;
  double dx, dy, dz, b0x, b0y, b0z, gamma;
  int limiter, heun, nxs;
  int wrap;  // bit a: axis a is periodic and the stage wraps its reads (its halo is not read)
  size_t off_ax, off_ay, off_az;
  // the face-flux record of a refined hierarchy's blocks (null: none): per block 6 slot
  // indices (face 2 axis + side, -1: not recorded), and per slot the 5 conserved fluxes
  // of each of the face's cells, face cell (u, v) in the two other axes in order
  const int* rec_slots = nullptr;
  double* rec = nullptr;
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

// Byte addressing of the state in the stage kernel: every array is a base pointer that
// is uniform over the workgroup (the array and the x-plane: scalar arithmetic) plus a
// per-thread 32-bit byte offset of the (y, z) column, so an element's address costs no
// vector multiply.  Planes are at most 4100 x 4101 doubles (the entry points' bound),
// far below 4 GiB.
struct Addr3 {
  size_t cplane, xplane, yplane, zplane;  // bytes per x-plane: a cell field, A_x, A_y, A_z
  size_t field;                           // bytes per cell field
  size_t ox, oy, oz;                      // byte offsets of A_x, A_y, A_z in the state
  uint32_t crow, arow;                    // bytes per y-row: cells and A_z; A_x and A_y
};

__device__ __forceinline__ Addr3 addr3(const StageArgs& a) {
  const Grid3 g = a.g;
  Addr3 d;
  d.crow = (uint32_t)g.Z * 8u;
  d.arow = (uint32_t)(g.Z + 1) * 8u;
  d.cplane = (size_t)g.Y * g.Z * 8;
  d.xplane = (size_t)(g.Y + 1) * (g.Z + 1) * 8;
  d.yplane = (size_t)g.Y * (g.Z + 1) * 8;
  d.zplane = (size_t)(g.Y + 1) * g.Z * 8;
  d.field = (size_t)g.X * d.cplane;
  d.ox = a.off_ax * 8;
  d.oy = a.off_ay * 8;
  d.oz = a.off_az * 8;
  return d;
}

__device__ __forceinline__ double ldb(const char* base, uint32_t off) {
  return *reinterpret_cast<const double*>(base + off);
}
__device__ __forceinline__ void stb(char* base, uint32_t off, double v) {
  *reinterpret_cast<double*>(base + off) = v;
}

// The primitive variables of one cell (mhd3d.py:475-509) from its 5 conserved values
// and the 12 potential values around it: ax[a][b] = A_x(i, j+a, k+b), ay[a][b] =
// A_y(i+a, j, k+b), az[a][b] = A_z(i+a, j+b, k).  Also the face fields of the cell's
// low x, y and z faces (fx, fy, fz, mean field included), which are the B_n of the
// faces the stage solves there.  FAST (the mixed path) multiplies by the reciprocal
// spacings and divides once for 1/rho; otherwise the divisions of the reference are
// kept for bit-for-bit parity (mhd3d.py:361-378).
template <bool FAST>
struct CellPrim {
  double r, u, v, w, p, bx, by, bz;
  double fx, fy, fz;
  __device__ __forceinline__ void finish(const StageArgs& a, const double (&ax)[2][2],
                                         const double (&ay)[2][2], const double (&az)[2][2],
                                         const double (&cv)[5], double rdx, double rdy,
                                         double rdz) {
    if (FAST) {
      const double bx0 = (az[0][1] - az[0][0]) * rdy - (ay[0][1] - ay[0][0]) * rdz;
      const double bx1 = (az[1][1] - az[1][0]) * rdy - (ay[1][1] - ay[1][0]) * rdz;
      const double by0 = (ax[0][1] - ax[0][0]) * rdz - (az[1][0] - az[0][0]) * rdx;
      const double by1 = (ax[1][1] - ax[1][0]) * rdz - (az[1][1] - az[0][1]) * rdx;
      const double bz0 = (ay[1][0] - ay[0][0]) * rdx - (ax[1][0] - ax[0][0]) * rdy;
      const double bz1 = (ay[1][1] - ay[0][1]) * rdx - (ax[1][1] - ax[0][1]) * rdy;
      bx = 0.5 * (bx0 + bx1) + a.b0x;
      by = 0.5 * (by0 + by1) + a.b0y;
      bz = 0.5 * (bz0 + bz1) + a.b0z;
      fx = bx0 + a.b0x;
      fy = by0 + a.b0y;
      fz = bz0 + a.b0z;
      r = cv[0];
      const double ri = 1.0 / r;
      u = cv[1] * ri;
      v = cv[2] * ri;
      w = cv[3] * ri;
    } else {
      const double bxa = (az[0][1] - az[0][0]) / a.dy - (ay[0][1] - ay[0][0]) / a.dz + a.b0x;
      const double bxb = (az[1][1] - az[1][0]) / a.dy - (ay[1][1] - ay[1][0]) / a.dz + a.b0x;
      const double bya = (ax[0][1] - ax[0][0]) / a.dz - (az[1][0] - az[0][0]) / a.dx + a.b0y;
      const double byb = (ax[1][1] - ax[1][0]) / a.dz - (az[1][1] - az[0][1]) / a.dx + a.b0y;
      const double bza = (ay[1][0] - ay[0][0]) / a.dx - (ax[1][0] - ax[0][0]) / a.dy + a.b0z;
      const double bzb = (ay[1][1] - ay[0][1]) / a.dx - (ax[1][1] - ax[0][1]) / a.dy + a.b0z;
      bx = 0.5 * (bxa + bxb);
      by = 0.5 * (bya + byb);
      bz = 0.5 * (bza + bzb);
      fx = bxa;
      fy = bya;
      fz = bza;
      r = cv[0];
      u = cv[1] / r;
      v = cv[2] / r;
      w = cv[3] / r;
    }
    p = (a.gamma - 1.0) * (cv[4] - 0.5 * r * (u * u + v * v + w * w) -
                           0.5 * (bx * bx + by * by + bz * bz));
  }
  // cell (i, j, k) of the haloed grid a.g, by element indices
  __device__ __forceinline__ void compute(const StageArgs& a, const EdgePot& A, size_t PL, int i,
                                          int j, int k, double rdx, double rdy, double rdz) {
    const size_t c = cidx(a.g, i, j, k);
    double ax[2][2], ay[2][2], az[2][2], cv[5];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        ax[s][t] = A.AX(i, j + s, k + t);
        ay[s][t] = A.AY(i + s, j, k + t);
        az[s][t] = A.AZ(i + s, j + t, k);
      }
#pragma unroll
    for (int q = 0; q < 5; ++q) cv[q] = a.in[q * PL + c];
    finish(a, ax, ay, az, cv, rdx, rdy, rdz);
  }
  // the cell of x-plane i whose column has the byte offsets c0 (cells, A_z) and c1
  // (A_x, A_y): the uniform plane bases carry every other displacement
  __device__ __forceinline__ void load(const StageArgs& a, const Addr3& d, const char* in, int i,
                                       uint32_t c0, uint32_t c1, double rdx, double rdy,
                                       double rdz) {
    const char* px = in + d.ox + (size_t)i * d.xplane;
    const char* py = in + d.oy + (size_t)i * d.yplane;
    const char* pz = in + d.oz + (size_t)i * d.zplane;
    const char* pc = in + (size_t)i * d.cplane;
    double ax[2][2], ay[2][2], az[2][2], cv[5];
    ax[0][0] = ldb(px, c1);
    ax[0][1] = ldb(px + 8, c1);
    ax[1][0] = ldb(px + d.arow, c1);
    ax[1][1] = ldb(px + d.arow + 8, c1);
    ay[0][0] = ldb(py, c1);
    ay[0][1] = ldb(py + 8, c1);
    ay[1][0] = ldb(py + d.yplane, c1);
    ay[1][1] = ldb(py + d.yplane + 8, c1);
    az[0][0] = ldb(pz, c0);
    az[0][1] = ldb(pz + d.crow, c0);
    az[1][0] = ldb(pz + d.zplane, c0);
    az[1][1] = ldb(pz + d.zplane + d.crow, c0);
#pragma unroll
    for (int q = 0; q < 5; ++q) cv[q] = ldb(pc + q * d.field, c0);
    finish(a, ax, ay, az, cv, rdx, rdy, rdz);
  }
  __device__ __forceinline__ bool bad() const {
    return !(finite(r) && finite(u) && finite(v) && finite(w) && finite(p) && finite(bx) &&
             finite(by) && finite(bz) && r > 0.0 && p > 0.0);
  }
};

// One stage of the tile (ty_, tz_) and x segment seg_ of the grid a.g (the tiles
// and segments of a launch together cover the grid).  LIM: the slope limiter fixed
// at compile time (the mixed path's branch-free FP32 flux), or -1 for the parity
// path's runtime limiter.  DIRECT (a block of a refined hierarchy): a cell outside the
// grid's interior is read in place from the neighbouring block of the same level
// holding it, `hood` giving the offsets (doubles from `root`) of the 27 blocks around
// this one (index (x + 1) * 9 + (y + 1) * 3 + (z + 1); -1: no such block, and the
// cell is read from this block's own ghosts, which must then be filled).
// `rslots` (a block of a refined hierarchy, or null): the block's 6 face slots in
// a.rec; the fluxes the stage solves on those faces of the block are written there as
// double, in the order of the cell fields (the refluxing reads them).
// EARLY (the EARLY_* bits): which stored values the CT and the update start from are
// loaded before the plane's barrier rather than where they are used -- the cells and
// the edges of `in`, and of `out` for the Heun average; early loads overlap their
// latency with the barrier and the ring plane's loads, at the price of the registers
// that hold them.  LDS_OUT_CELLS: the 5 cell values of `out` go to LDS instead
// (`dma`: 10 x THREADS words, the low and high halves of each), by loads that write
// LDS directly and hold no register.
constexpr int EARLY_IN_CELLS = 1, EARLY_OUT_CELLS = 2, LDS_OUT_CELLS = 4, EARLY_IN_EDGES = 8,
              EARLY_OUT_EDGES = 16;
constexpr int EARLY_ALL = EARLY_IN_CELLS | EARLY_OUT_CELLS | EARLY_IN_EDGES | EARLY_OUT_EDGES;
template <typename P, int TY, int TZ, int LIM, bool DIRECT = false, int EARLY = EARLY_ALL>
__device__ __forceinline__ void stage3d_body(const StageArgs& a, int ty_, int tz_, int seg_,
                                             StageShared<P, TY, TZ>& sh,
                                             const long long* hood = nullptr,
                                             const double* root = nullptr,
                                             const int* rslots = nullptr,
                                             uint32_t* dma = nullptr) {
  using T = Tile<TY, TZ>;
  constexpr bool FAST = sizeof(P) == 4;
  const Grid3 g = a.g;
  const int tid = threadIdx.x;
  const int jy0 = 2 + ty_ * TY, kz0 = 2 + tz_ * TZ;       // first interior cell of the tile
  const int ix0 = 2 + seg_ * a.nxs;                       // first plane of the segment
  const int tyv = min(TY, g.ny + 2 - jy0), tzv = min(TZ, g.nz + 2 - kz0);
  const int nxv = min(a.nxs, g.nx + 2 - ix0);
  const bool last_y = jy0 + TY >= g.ny + 2, last_z = kz0 + TZ >= g.nz + 2;
  const bool last_x = ix0 + a.nxs >= g.nx + 2;
  // this thread's face column (local y, z in [-1, T])
  const bool fthread = tid < T::NF;
  const int fa = tid / T::FZ, fb = tid % T::FZ;
  const int ly = fa - 1, lz = fb - 1;
  const int gy = jy0 + ly, gz = kz0 + lz;
  const int cgy = min(max(gy, 0), g.Y - 1), cgz = min(max(gz, 0), g.Z - 1);
  // the cell a read of haloed index `h` along a y or z axis takes: its periodic image
  // among the interior cells where the stage wraps that axis, else the clamped index
  // (the halo, filled before the launch)
  auto read_index = [&](int h, int n, int extent, int bit) {
    if (a.wrap & bit) {
      int t = (h - 2) % n;
      return (t < 0 ? t + n : t) + 2;
    }
    return min(max(h, 0), extent - 1);
  };
  // DIRECT: the block (-1, 0, +1) holding haloed index h along an axis of n interior
  // cells, and h in that block's haloed indices
  auto rel = [](int h, int n) { return h < 2 ? -1 : (h >= n + 2 ? 1 : 0); };
  const int yr = DIRECT ? rel(gy, g.ny) : 0, zr = DIRECT ? rel(gz, g.nz) : 0;
  const int ry = read_index(gy, g.ny, g.Y, 2), rz = read_index(gz, g.nz, g.Z, 4);
  const Addr3 d = addr3(a);
  const char* inb = reinterpret_cast<const char*>(a.in);
  char* outb = reinterpret_cast<char*>(a.out);
  // byte offsets of the column: c0/c1 where the column writes (cells and A_z / A_x and
  // A_y), w0/w1 where it reads its cells
  const uint32_t c0 = (uint32_t)cgy * d.crow + (uint32_t)cgz * 8u;
  const uint32_t c1 = (uint32_t)cgy * d.arow + (uint32_t)cgz * 8u;
  const uint32_t w0 = (uint32_t)ry * d.crow + (uint32_t)rz * 8u;
  const uint32_t w1 = (uint32_t)ry * d.arow + (uint32_t)rz * 8u;
  // DIRECT: the column's offsets in the neighbouring block holding it
  const int ly_ = gy - yr * g.ny, lz_ = gz - zr * g.nz;
  const uint32_t n0 = (uint32_t)ly_ * d.crow + (uint32_t)lz_ * 8u;
  const uint32_t n1 = (uint32_t)ly_ * d.arow + (uint32_t)lz_ * 8u;
  // the x-plane a read of plane `pl` (in [ix0 - 2, ix0 + nxv + 1]) takes
  auto read_plane = [&](int pl) {
    if (a.wrap & 1) return pl < 2 ? pl + g.nx : (pl >= g.nx + 2 ? pl - g.nx : pl);
    return min(max(pl, 0), g.X - 1);
  };
  // the state a read of plane `pl` of a column in block (., yrel, zrel) takes, the
  // plane there and the column's offsets there (own: h0/h1, in the neighbour: l0/l1)
  auto source = [&](int pl, int yrel, int zrel, uint32_t h0, uint32_t h1, uint32_t l0,
                    uint32_t l1, int& lp, uint32_t& o0, uint32_t& o1) -> const char* {
    if constexpr (DIRECT) {
      const int xr = rel(pl, g.nx);
      const long long off = hood[(xr + 1) * 9 + (yrel + 1) * 3 + zrel + 1];
      if (off >= 0) {
        lp = pl - xr * g.nx;
        o0 = l0;
        o1 = l1;
        return reinterpret_cast<const char*>(root + off);
      }
    }
    lp = read_plane(pl);
    o0 = h0;
    o1 = h1;
    return inb;
  };
  const double dt = *a.dt;
  const double dtdx = dt / a.dx, dtdy = dt / a.dy, dtdz = dt / a.dz;
  const double rdx = 1.0 / a.dx, rdy = 1.0 / a.dy, rdz = 1.0 / a.dz;
  const P gam = P(a.gamma), inv_gm1 = P(1.0 / (a.gamma - 1.0));
  // the block's 6 record slots (-1: none; uniform over the workgroup)
  int slot[6] = {-1, -1, -1, -1, -1, -1};
  if (rslots) {
#pragma unroll
    for (int s6 = 0; s6 < 6; ++s6) slot[s6] = rslots[s6];
  }
  // slot `sl`, face cell `cell` (square blocks: g.ny * g.nz cells per face): the 5 fluxes
  // of the cell fields, which sit in rows row[0..4] of the face's frame
  auto record = [&](int sl, int cell, const P* fl, const int (&row)[5]) {
    double* o = a.rec + ((size_t)sl * (g.ny * g.nz) + cell) * 5;
#pragma unroll
    for (int r = 0; r < 5; ++r) o[r] = double(fl[row[r]]);
  };

  // The mixed path's x-sweep computes each cell's slope once: the right state of face f
  // and the left state of face f+1 (carried in lcar) come from the slope of cell f, so
  // its ring needs planes f-1 .. f+1 only; the parity path keeps the reference's
  // per-face reconstruction from planes f-2 .. f+1.
  constexpr bool CARRY = sizeof(P) == 4 && LIM >= 0;
  constexpr int NQ = CARRY ? 3 : 4;
  // register state of the column
  P q[7][NQ];                      // x-frame primitives of the last NQ planes (to f+1)
  P lcar[7];                       // CARRY: the left state of x-face f+1
  P own_bx = P(0);                 // cell B_x of plane f+1 (the ring holds the rest), for LDS
  P bnp[3], bnn[3];                // face fields of the column's low x/y/z faces: planes f, f+1
  P fprev[5], fcur[5];             // x-face fluxes (rows 0-4) of faces f-1 and f
  // dt times the y/z flux differences of the cell of plane f-1, for its update; the
  // mixed path keeps their sum (the parity path adds them in the reference's order)
  double dyy[5], dzz[FAST ? 1 : 5];
#pragma unroll
  for (int c = 0; c < 7; ++c) {
    lcar[c] = P(0);
#pragma unroll
    for (int s = 0; s < NQ; ++s) q[c][s] = P(0);
  }
#pragma unroll
  for (int r = 0; r < 5; ++r) { fprev[r] = fcur[r] = P(0); dyy[r] = 0.0; }
#pragma unroll
  for (int r = 0; r < (FAST ? 1 : 5); ++r) dzz[r] = 0.0;
#pragma unroll
  for (int r = 0; r < 3; ++r) bnp[r] = bnn[r] = P(0);

  // the own column of plane `pl`: registers (ring + own), cell E, bad cells
  auto own_column = [&](int pl) {
    if (!fthread) return;
    CellPrim<FAST> c;
    int lp;
    uint32_t o0, o1;
    const char* src = source(pl, yr, zr, w0, w1, n0, n1, lp, o0, o1);
    c.load(a, d, src, lp, o0, o1, rdx, rdy, rdz);
    bnp[0] = bnn[0]; bnp[1] = bnn[1]; bnp[2] = bnn[2];
    bnn[0] = P(c.fx); bnn[1] = P(c.fy); bnn[2] = P(c.fz);
    const P x[7] = {P(c.r), P(c.u), P(c.v), P(c.w), P(c.p), P(c.by), P(c.bz)};
    own_bx = P(c.bx);
#pragma unroll
    for (int k = 0; k < 7; ++k) {
#pragma unroll
      for (int s = 0; s + 1 < NQ; ++s) q[k][s] = q[k][s + 1];
      q[k][NQ - 1] = x[k];
    }
    sh.ex[pl & 1][tid] = P(c.by * c.w - c.bz * c.v);
    sh.ey[pl % 3][tid] = P(c.bz * c.u - c.bx * c.w);
    sh.ez[pl % 3][tid] = P(c.bx * c.v - c.by * c.u);
    if (ly >= 0 && ly < tyv && lz >= 0 && lz < tzv && pl >= ix0 && pl < ix0 + nxv && c.bad())
      atomicAdd(a.bad, 1);
  };
  // the outermost ring of primitive columns but its four corners, which no face stencil
  // reads (the y-faces read columns 1..PZ-2, the z-faces rows 1..PY-2): one column per
  // thread (tid < NH; 64 for the 12 x 16 tile, one wave)
  constexpr int NH = 2 * (T::PZ - 2) + 2 * (T::PY - 2);
  static_assert(NH <= T::THREADS, "one ring column per thread");
  const bool ring = tid < NH;
  int rpy, rpz;
  if (tid < T::PZ - 2) { rpy = 0; rpz = tid + 1; }
  else if (tid < 2 * (T::PZ - 2)) { rpy = T::PY - 1; rpz = tid - (T::PZ - 2) + 1; }
  else { const int e = tid - 2 * (T::PZ - 2); rpy = 1 + e / 2; rpz = (e & 1) ? T::PZ - 1 : 0; }
  const int hj = jy0 + rpy - 2, hk = kz0 + rpz - 2;  // the ring column's haloed indices
  const int ryr = DIRECT ? rel(hj, g.ny) : 0, rzr = DIRECT ? rel(hk, g.nz) : 0;
  const int rj = read_index(hj, g.ny, g.Y, 2), rk = read_index(hk, g.nz, g.Z, 4);
  const uint32_t r0 = (uint32_t)rj * d.crow + (uint32_t)rk * 8u;
  const uint32_t r1 = (uint32_t)rj * d.arow + (uint32_t)rk * 8u;
  const int lj_ = hj - ryr * g.ny, lk_ = hk - rzr * g.nz;
  const uint32_t m0 = (uint32_t)lj_ * d.crow + (uint32_t)lk_ * 8u;
  const uint32_t m1 = (uint32_t)lj_ * d.arow + (uint32_t)lk_ * 8u;
  // plane `pl` into LDS: the own column from registers, the outer ring computed here
  auto store_plane = [&](int pl) {
    if (fthread) {
      const int idx = (fa + 1) * T::PZ + (fb + 1);
#pragma unroll
      for (int r = 0; r < 5; ++r) sh.prim[r][idx] = q[r][NQ - 1];
      sh.prim[5][idx] = own_bx;
      sh.prim[6][idx] = q[5][NQ - 1];
      sh.prim[7][idx] = q[6][NQ - 1];
    }
    if (ring) {
      CellPrim<FAST> c;
      int lp;
      uint32_t o0, o1;
      const char* src = source(pl, ryr, rzr, r0, r1, m0, m1, lp, o0, o1);
      c.load(a, d, src, lp, o0, o1, rdx, rdy, rdz);
      const int idx = rpy * T::PZ + rpz;
      sh.prim[0][idx] = P(c.r); sh.prim[1][idx] = P(c.u); sh.prim[2][idx] = P(c.v);
      sh.prim[3][idx] = P(c.w); sh.prim[4][idx] = P(c.p); sh.prim[5][idx] = P(c.bx);
      sh.prim[6][idx] = P(c.by); sh.prim[7][idx] = P(c.bz);
    }
  };

  // prologue: planes ix0-2 (ring) and ix0-1 (ring + LDS)
  own_column(ix0 - 2);
  own_column(ix0 - 1);
  store_plane(ix0 - 1);
  __syncthreads();

  for (int f = ix0 - 1; f <= ix0 + nxv; ++f) {
    const int cur = f & 1, prv = cur ^ 1;
    if (fthread) {
      // a. y- and z-faces of plane f below the column's cell, from the LDS plane
      //    (their B_n: the column's face fields of plane f)
      const int pidx = (fa + 1) * T::PZ + (fb + 1);
      if (fa >= 1) {
        constexpr int comp[7] = {0, 2, 3, 1, 4, 7, 5};
        P qq[7][4];
#pragma unroll
        for (int c = 0; c < 7; ++c)
#pragma unroll
          for (int s = 0; s < 4; ++s) qq[c][s] = sh.prim[comp[c]][pidx + (s - 2) * T::PZ];
        P fl[7];
        face_solve<P, LIM>(qq, a.limiter, bnn[1], gam, inv_gm1, fl);
#pragma unroll
        for (int r = 1; r < 6; ++r) sh.fy[r - 1][tid] = fl[r];
        sh.fyc[cur][0][tid] = fl[0];
        sh.fyc[cur][1][tid] = fl[6];
        if (rslots && f >= ix0 && f < ix0 + nxv && lz >= 0 && lz < tzv) {
          const int sl = gy == 2 ? slot[2] : (gy == g.ny + 2 ? slot[3] : -1);
          constexpr int yrow[5] = {0, 3, 1, 2, 4};  // y-frame rows of rho, m_x, m_y, m_z, E
          if (sl >= 0) record(sl, (f - 2) * g.nz + (gz - 2), fl, yrow);
        }
      }
      if (fb >= 1) {
        constexpr int comp[7] = {0, 3, 1, 2, 4, 5, 6};
        P qq[7][4];
#pragma unroll
        for (int c = 0; c < 7; ++c)
#pragma unroll
          for (int s = 0; s < 4; ++s) qq[c][s] = sh.prim[comp[c]][pidx + (s - 2)];
        P fl[7];
        face_solve<P, LIM>(qq, a.limiter, bnn[2], gam, inv_gm1, fl);
#pragma unroll
        for (int r = 1; r < 5; ++r) sh.fz[r - 1][tid] = fl[r];
        sh.fz[4][tid] = fl[6];
        sh.fzc[cur][0][tid] = fl[0];
        sh.fzc[cur][1][tid] = fl[5];
        if (rslots && f >= ix0 && f < ix0 + nxv && ly >= 0 && ly < tyv) {
          const int sl = gz == 2 ? slot[4] : (gz == g.nz + 2 ? slot[5] : -1);
          constexpr int zrow[5] = {0, 2, 3, 1, 4};  // z-frame rows of rho, m_x, m_y, m_z, E
          if (sl >= 0) record(sl, (f - 2) * g.ny + (gy - 2), fl, zrow);
        }
      }
    }
    // b. the own column of plane f+1 into the ring; c. x-face f from the ring
    own_column(f + 1);
    auto keep_x_face = [&](const P* fl) {
#pragma unroll
      for (int r = 0; r < 5; ++r) {
        fprev[r] = fcur[r];
        fcur[r] = fl[r];
      }
      sh.fx[0][tid] = fl[0];
      sh.fx[1][tid] = fl[5];
      sh.fx[2][tid] = fl[6];
      const int sxf = f == 2 ? slot[0] : (f == g.nx + 2 ? slot[1] : -1);
      if (sxf >= 0 && ly >= 0 && ly < tyv && lz >= 0 && lz < tzv) {
        constexpr int xrow[5] = {0, 1, 2, 3, 4};
        record(sxf, (gy - 2) * g.nz + (gz - 2), fl, xrow);
      }
    };
    if constexpr (CARRY) {
      if (fthread) {
        // the slope of cell f (ring: planes f-1, f, f+1; mhd_sweep_pyccel.py:290-307)
        P rs[7], ln[7];
#pragma unroll
        for (int c = 0; c < 7; ++c) {
          const P sl = fast::slope<(CARRY ? LIM : 0)>(q[c][1] - q[c][0], q[c][2] - q[c][1]);
          rs[c] = q[c][1] - 0.5f * sl;
          ln[c] = q[c][1] + 0.5f * sl;
        }
        if (f >= ix0) {
          const bool first = lcar[0] <= 0.0f || lcar[4] <= 0.0f || rs[0] <= 0.0f || rs[4] <= 0.0f;
          P ls[7];
#pragma unroll
          for (int c = 0; c < 7; ++c) {
            ls[c] = fast::sel(first, q[c][0], lcar[c]);
            rs[c] = fast::sel(first, q[c][1], rs[c]);
          }
          P fl[7];
          fast::hlld(ls, rs, bnp[0], gam, inv_gm1, fl);
          keep_x_face(fl);
        }
#pragma unroll
        for (int c = 0; c < 7; ++c) lcar[c] = ln[c];
      }
    } else if (fthread && f >= ix0) {
      P fl[7];
      face_solve<P, LIM>(q, a.limiter, bnp[0], gam, inv_gm1, fl);
      keep_x_face(fl);
    }
    // d. the stored values the CT and the update below start from (the potentials of
    //    plane f's edges, the cells of plane f-1; with the Heun average also their
    //    values in `out`), loaded before the barrier so that their latency overlaps it
    //    and the loads of the ring plane
    const bool in_y = ly >= 0 && (ly < tyv || (ly == tyv && last_y));
    const bool in_z = lz >= 0 && (lz < tzv || (lz == tzv && last_z));
    const bool cell_y = ly >= 0 && ly < tyv, cell_z = lz >= 0 && lz < tzv;
    const bool node_x = f >= ix0 && (f < ix0 + nxv || (f == ix0 + nxv && last_x));
    const bool do_ex = fthread && f >= ix0 && f < ix0 + nxv && in_y && in_z;
    const bool do_ey = fthread && node_x && cell_y && in_z;
    const bool do_ez = fthread && node_x && in_y && cell_z;
    const bool do_cell = fthread && cell_y && cell_z && f - 1 >= ix0 && f - 1 < ix0 + nxv;
    const size_t oex = d.ox + (size_t)f * d.xplane, oey = d.oy + (size_t)f * d.yplane;
    const size_t oez = d.oz + (size_t)f * d.zplane, ocell = (size_t)(f - 1) * d.cplane;
    double vin[8], vout[8];  // A_x, A_y, A_z at the edges, then the 5 cell fields
#pragma unroll
    for (int r = 0; r < 8; ++r) vin[r] = vout[r] = 0.0;
    constexpr bool in_e = EARLY & EARLY_IN_EDGES, out_e = EARLY & EARLY_OUT_EDGES;
    constexpr bool in_c = EARLY & EARLY_IN_CELLS;
    constexpr bool out_c = EARLY & (EARLY_OUT_CELLS | LDS_OUT_CELLS);
    if constexpr (in_e) {
      if (do_ex) vin[0] = ldb(inb + oex, c1);
      if (do_ey) vin[1] = ldb(inb + oey, c1);
      if (do_ez) vin[2] = ldb(inb + oez, c0);
    }
    if constexpr (in_c) {
      if (do_cell) {
#pragma unroll
        for (int r = 0; r < 5; ++r) vin[3 + r] = ldb(inb + ocell + r * d.field, c0);
      }
    }
    if (out_e && a.heun) {
      if (do_ex) vout[0] = ldb(outb + oex, c1);
      if (do_ey) vout[1] = ldb(outb + oey, c1);
      if (do_ez) vout[2] = ldb(outb + oez, c0);
    }
    if ((EARLY & EARLY_OUT_CELLS) && a.heun) {
      if (do_cell) {
#pragma unroll
        for (int r = 0; r < 5; ++r) vout[3 + r] = ldb(outb + ocell + r * d.field, c0);
      }
    }
    if constexpr ((EARLY & LDS_OUT_CELLS) != 0) {
      // every face column's cell of plane f - 1 (in the block: plane >= ix0 - 2 >= 0,
      // clamped column), whether or not it is updated; lane l of wave w writes word
      // w * 64 + l of each row
      if (a.heun && fthread) {
        const char* src = outb + ocell + c0;
        uint32_t* row = dma + (tid & ~63);
#pragma unroll
        for (int r = 0; r < 5; ++r)
#pragma unroll
          for (int h = 0; h < 2; ++h)
            __builtin_amdgcn_global_load_lds(
                (const __attribute__((address_space(1))) void*)(src + r * d.field + 4 * h),
                (__attribute__((address_space(3))) void*)(row + (2 * r + h) * T::THREADS), 4, 0,
                0);
      }
    }
    __syncthreads();
    // e. the LDS plane is free: plane f+1 goes in for the next iteration
    store_plane(f + 1);
    if constexpr ((EARLY & LDS_OUT_CELLS) != 0) {
      // the ring plane's loads were issued after these and have been waited for, so
      // the wait for every vector memory operation here costs nothing more
      if (a.heun && do_cell) {
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
#pragma unroll
        for (int r = 0; r < 5; ++r)
          vout[3 + r] = __hiloint2double((int)dma[(2 * r + 1) * T::THREADS + tid],
                                         (int)dma[2 * r * T::THREADS + tid]);
      }
    }
    if (fthread) {
      const int e0 = (f + 2) % 3, e1 = f % 3;  // planes f-1 and f of the E rings
      const int tym = tid - T::FZ, tzm = tid - 1, tymz = tid - T::FZ - 1;
      // f1. E_x on the x-cell plane f at (y-node ly, z-node lz): plane (a=y, b=z)
      if (do_ex) {
        const double e = ct_node(
            double(sh.fy[4][tzm]), double(sh.fy[4][tid]),
            double(sh.fz[4][tym]), double(sh.fz[4][tid]),
            double(sh.fyc[cur][0][tzm]), double(sh.fyc[cur][0][tid]),
            double(sh.fzc[cur][0][tym]), double(sh.fzc[cur][0][tid]),
            double(sh.ex[f & 1][tymz]), double(sh.ex[f & 1][tzm]), double(sh.ex[f & 1][tym]),
            double(sh.ex[f & 1][tid]));
        // E_x's A_x at (f, gy, gz): column c1
        double v = (in_e ? vin[0] : ldb(inb + oex, c1)) - dt * e;
        if (a.heun) v = (v + (out_e ? vout[0] : ldb(outb + oex, c1))) * 0.5;
        stb(outb + oex, c1, v);
      }
      // f2. E_y at (x-node f, y-cell ly, z-node lz): plane (a=z, b=x)
      if (do_ey) {
        const double e = ct_node(
            double(sh.fzc[prv][1][tid]), double(sh.fzc[cur][1][tid]),
            double(sh.fx[2][tzm]), double(sh.fx[2][tid]),
            double(sh.fzc[prv][0][tid]), double(sh.fzc[cur][0][tid]),
            double(sh.fx[0][tzm]), double(sh.fx[0][tid]),
            double(sh.ey[e0][tzm]), double(sh.ey[e0][tid]), double(sh.ey[e1][tzm]),
            double(sh.ey[e1][tid]));
        double v = (in_e ? vin[1] : ldb(inb + oey, c1)) - dt * e;
        if (a.heun) v = (v + (out_e ? vout[1] : ldb(outb + oey, c1))) * 0.5;
        stb(outb + oey, c1, v);
      }
      // f3. E_z at (x-node f, y-node ly, z-cell lz): plane (a=x, b=y)
      if (do_ez) {
        const double e = ct_node(
            double(sh.fx[1][tym]), double(sh.fx[1][tid]),
            double(sh.fyc[prv][1][tid]), double(sh.fyc[cur][1][tid]),
            double(sh.fx[0][tym]), double(sh.fx[0][tid]),
            double(sh.fyc[prv][0][tid]), double(sh.fyc[cur][0][tid]),
            double(sh.ez[e0][tym]), double(sh.ez[e1][tym]), double(sh.ez[e0][tid]),
            double(sh.ez[e1][tid]));
        double v = (in_e ? vin[2] : ldb(inb + oez, c0)) - dt * e;
        if (a.heun) v = (v + (out_e ? vout[2] : ldb(outb + oez, c0))) * 0.5;
        stb(outb + oez, c0, v);
      }
      // g. update of cell (f-1, ly, lz): x-faces f-1 and f, y/z terms of plane f-1
      if (cell_y && cell_z) {
        if (do_cell) {
#pragma unroll
          for (int r = 0; r < 5; ++r) {
            const double dfx = double(fcur[r]) - double(fprev[r]);
            const size_t o = ocell + r * d.field;
            const double u0 = in_c ? vin[3 + r] : ldb(inb + o, c0);
            double v;
            if constexpr (FAST) v = u0 - (dtdx * dfx + dyy[r]);
            else v = u0 - (dtdx * dfx + dyy[r] + dzz[r]);
            if (a.heun) v = (v + (out_c ? vout[3 + r] : ldb(outb + o, c0))) * 0.5;
            stb(outb + o, c0, v);
          }
        }
        // y/z flux differences of cell (f, ly, lz) for the next iteration
        if (f >= ix0 && f < ix0 + nxv) {
          const int typ = tid + T::FZ, tzp = tid + 1;
          // rows of the y-sweep (rho, m_y, m_z, m_x, E) and z-sweep (rho, m_z, m_x, m_y, E);
          // row 0 (mass) sits in fyc/fzc, rows r >= 1 at r - 1
          constexpr int yrow[5] = {0, 3, 1, 2, 4};
          constexpr int zrow[5] = {0, 2, 3, 1, 4};
#pragma unroll
          for (int r = 0; r < 5; ++r) {
            const P* fyr = r == 0 ? sh.fyc[cur][0] : sh.fy[yrow[r] - 1];
            const P* fzr = r == 0 ? sh.fzc[cur][0] : sh.fz[zrow[r] - 1];
            const double ty = dtdy * (double(fyr[typ]) - double(fyr[tid]));
            const double tz = dtdz * (double(fzr[tzp]) - double(fzr[tid]));
            if constexpr (FAST) {
              dyy[r] = ty + tz;
            } else {
              dyy[r] = ty;
              dzz[r] = tz;
            }
          }
        }
      }
    }
    __syncthreads();
  }
}

// The stage over a uniform grid, one workgroup per (tile, segment).  Workgroup
// order: the hardware hands workgroup w to XCD w % 8, each XCD with its own L2.
// The workgroups are renumbered so that XCD x runs a contiguous run of tiles (z
// fastest, then y, then the x segment): tiles running together on one XCD are
// neighbours, and the halo lines they share are fetched into that L2 once.
template <typename P, int TY, int TZ, int WAVES, int LIM, int EARLY>
__global__ void __launch_bounds__((Tile<TY, TZ>::THREADS), WAVES)
k_stage3d(StageArgs a) {
  __shared__ StageShared<P, TY, TZ> sh;
  __shared__ uint32_t dma[(EARLY & LDS_OUT_CELLS) ? 10 * Tile<TY, TZ>::THREADS : 1];
  const Grid3 g = a.g;
  const int tiles_z = (g.nz + TZ - 1) / TZ, tiles_y = (g.ny + TY - 1) / TY;
  const int n = gridDim.x, xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
  const int per = n >> 3, rem = n & 7;
  const int logical = xcd * per + min(xcd, rem) + k;
  const int tz = logical % tiles_z;
  const int rest = logical / tiles_z;
  stage3d_body<P, TY, TZ, LIM, false, EARLY>(a, rest % tiles_y, tz, rest / tiles_y, sh, nullptr,
                                            nullptr, nullptr, dma);
}


// WAVES: the waves per SIMD the register allocation must allow (2 or 3 for the
// mixed path: 3 trades a few spills for occupancy; the autotuner picks).
template <typename P, int TY, int TZ, int WAVES, int LIM = -1, int EARLY = EARLY_ALL>
inline int launch_stage3d(const StageArgs& a, hipStream_t s) {
  using T = Tile<TY, TZ>;
  const int tiles = ((a.g.ny + TY - 1) / TY) * ((a.g.nz + TZ - 1) / TZ);
  const int segs = (a.g.nx + a.nxs - 1) / a.nxs;
  hipLaunchKernelGGL((k_stage3d<P, TY, TZ, WAVES, LIM, EARLY>), dim3(tiles * segs), dim3(T::THREADS),
                     0, s, a);
  return wb_last_error();
}

}  // namespace wb
