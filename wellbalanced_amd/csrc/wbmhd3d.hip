// 3D ideal-MHD kernels for gfx950, one kernel per stage piece: the halo fill,
// primitive variables with the face field from the edge potential, the MUSCL+HLLD
// sweep along each axis, the CT-contact edge field, the stage update with the
// fused Heun average, and the CFL reduction.  FP64, built with -ffp-contract=off.
//
// Layout: C order, z contiguous.  A haloed block is X = nx+4, Y = ny+4, Z = nz+4
// cells.  The state is one buffer: 5 cell planes (rho, mx, my, mz, E), then the
// edge potential A_x (X, Y+1, Z+1), A_y (X+1, Y, Z+1), A_z (X+1, Y+1, Z).
#include "mhd_device.hpp"
#include "rate3d.hpp"
#include "wbmhd_common.hpp"

using namespace wb;

namespace {

struct Dims3 {
  int X, Y, Z;
  __host__ __device__ size_t cells() const { return (size_t)X * Y * Z; }
};

// Extent of array `a` of the state (0..4 cells, 5..7 = A_x, A_y, A_z) along axis d.
__host__ __device__ inline int extent(const Dims3& g, int a, int d) {
  const int n = d == 0 ? g.X : (d == 1 ? g.Y : g.Z);
  if (a < 5) return n;
  return (a - 5 == d) ? n : n + 1;
}

__host__ __device__ inline size_t array_offset(const Dims3& g, int a) {
  if (a < 5) return (size_t)a * g.cells();
  size_t off = 5 * g.cells();
  for (int b = 5; b < a; ++b)
    off += (size_t)extent(g, b, 0) * extent(g, b, 1) * extent(g, b, 2);
  return off;
}

struct Pot {
  const double *ax, *ay, *az;
  int X, Y, Z;
  __device__ double AX(int i, int j, int k) const { return ax[((size_t)i * (Y + 1) + j) * (Z + 1) + k]; }
  __device__ double AY(int i, int j, int k) const { return ay[((size_t)i * Y + j) * (Z + 1) + k]; }
  __device__ double AZ(int i, int j, int k) const { return az[((size_t)i * (Y + 1) + j) * Z + k]; }
  // B_a = B0_a + d_b A_c - d_c A_b on face (i,j,k) normal to a (mhd3d.py:361-378).
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

__device__ inline Pot pot_of(const double* state, const Dims3& g) {
  Pot p;
  p.ax = state + array_offset(g, 5);
  p.ay = state + array_offset(g, 6);
  p.az = state + array_offset(g, 7);
  p.X = g.X;
  p.Y = g.Y;
  p.Z = g.Z;
  return p;
}

// mhd3d.py:475-509: prim (8, X, Y, Z) of the haloed block and the three face
// fields; counts the inner cells that are not finite or not positive.
__global__ void k_primitive3d(const double* __restrict__ state, Dims3 g, double dx, double dy,
                              double dz, double b0x, double b0y, double b0z, double gamma,
                              double* __restrict__ prim, double* __restrict__ fbx,
                              double* __restrict__ fby, double* __restrict__ fbz,
                              int* __restrict__ bad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  const int j = t / (g.Z + 1), k = t % (g.Z + 1);
  if (j > g.Y || i > g.X) return;
  const Pot A = pot_of(state, g);
  const int X = g.X, Y = g.Y, Z = g.Z;
  if (j < Y && k < Z) fbx[((size_t)i * Y + j) * Z + k] = A.bx(i, j, k, dy, dz, b0x);
  if (i < X && k < Z) fby[((size_t)i * (Y + 1) + j) * Z + k] = A.by(i, j, k, dz, dx, b0y);
  if (i < X && j < Y) fbz[((size_t)i * Y + j) * (Z + 1) + k] = A.bz(i, j, k, dx, dy, b0z);
  if (i >= X || j >= Y || k >= Z) return;
  const size_t c = ((size_t)i * Y + j) * Z + k;
  const size_t P = g.cells();
  const double bx = 0.5 * (A.bx(i, j, k, dy, dz, b0x) + A.bx(i + 1, j, k, dy, dz, b0x));
  const double by = 0.5 * (A.by(i, j, k, dz, dx, b0y) + A.by(i, j + 1, k, dz, dx, b0y));
  const double bz = 0.5 * (A.bz(i, j, k, dx, dy, b0z) + A.bz(i, j, k + 1, dx, dy, b0z));
  const double r = state[c];
  const double u = state[P + c] / r, v = state[2 * P + c] / r, w = state[3 * P + c] / r;
  const double p = (gamma - 1.0) * (state[4 * P + c] - 0.5 * r * (u * u + v * v + w * w) -
                                    0.5 * (bx * bx + by * by + bz * bz));
  prim[c] = r;
  prim[P + c] = u;
  prim[2 * P + c] = v;
  prim[3 * P + c] = w;
  prim[4 * P + c] = p;
  prim[5 * P + c] = bx;
  prim[6 * P + c] = by;
  prim[7 * P + c] = bz;
  const int h = HALO;
  if (i >= h && i < X - h && j >= h && j < Y - h && k >= h && k < Z - h) {
    const bool ok = finite(r) && finite(u) && finite(v) && finite(w) && finite(p) && finite(bx) &&
                    finite(by) && finite(bz) && r > 0.0 && p > 0.0;
    if (!ok) atomicAdd(bad, 1);
  }
}

// mhd3d.py:511-528 + mhd_sweep_pyccel.py:254-319: the face fluxes along S of
// every transverse cell, in the physical layout (7, ...) with n_S - 3 faces along
// S.  Rows: rho, m_S, m_b, m_c, E, B_b, B_c for the cyclic frame (S, b, c).
template <int S>
__global__ void __launch_bounds__(128) k_sweep3d(const double* __restrict__ prim, const double* __restrict__ bn, Dims3 g,
                          int limiter, double gamma, double* __restrict__ flux) {
  constexpr int B = (S + 1) % 3, C = (S + 2) % 3;
  constexpr int comp[7] = {0, 1 + S, 1 + B, 1 + C, 4, 5 + B, 5 + C};
  int n[3] = {g.X, g.Y, g.Z};
  int m[3] = {g.X, g.Y, g.Z};
  m[S] -= 3;  // faces along S
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i0 = blockIdx.y;
  const int i2 = t % m[2], i1 = t / m[2];
  if (i1 >= m[1] || i0 >= m[0]) return;
  int p[3] = {i0, i1, i2};
  const size_t P = g.cells();
  size_t stride[3] = {(size_t)n[1] * n[2], (size_t)n[2], 1};
  size_t base = p[0] * stride[0] + p[1] * stride[1] + p[2] * stride[2];
  double q[7][4];
#pragma unroll
  for (int c = 0; c < 7; ++c)
#pragma unroll
    for (int s = 0; s < 4; ++s) q[c][s] = prim[comp[c] * P + base + s * stride[S]];
  // face field: n_S + 1 along S, face f + 2
  int nb[3] = {n[0], n[1], n[2]};
  nb[S] += 1;
  int pb[3] = {p[0], p[1], p[2]};
  pb[S] += 2;
  const double normal = bn[((size_t)pb[0] * nb[1] + pb[1]) * nb[2] + pb[2]];
  double L[7], R[7], f[7];
  reconstruct<double>(q, limiter, L, R);
  hlld<double>(L, R, normal, gamma, f);
  const size_t FS = (size_t)m[0] * m[1] * m[2];
  const size_t o = ((size_t)p[0] * m[1] + p[1]) * m[2] + p[2];
#pragma unroll
  for (int c = 0; c < 7; ++c) flux[c * FS + o] = f[c];
}

// mhd3d.py:530-559: E along C on the block's C-edges by the CT-contact average
// of the A-sweep's flux of B_B (-E_C) and the B-sweep's flux of B_A (+E_C),
// (A, B) = (C+1, C+2).  Output in the physical layout: n_C along C, n+1 along A, B.
template <int CC>
__global__ void k_edge3d(const double* __restrict__ prim, const double* __restrict__ fa,
                         const double* __restrict__ fb, Dims3 g, double* __restrict__ e) {
  constexpr int A = (CC + 1) % 3, B = (CC + 2) % 3;
  const int n[3] = {g.X - 4, g.Y - 4, g.Z - 4};
  int m[3] = {n[0] + 1, n[1] + 1, n[2] + 1};
  m[CC] = n[CC];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int o0 = blockIdx.y;
  const int o2 = t % m[2], o1 = t / m[2];
  if (o1 >= m[1] || o0 >= m[0]) return;
  const int o[3] = {o0, o1, o2};
  const int kc = o[CC] + HALO, i = o[A] + 2, j = o[B] + 2;
  const int N[3] = {g.X, g.Y, g.Z};
  // flux_A: N_A - 3 faces along A
  int ma[3] = {N[0], N[1], N[2]};
  ma[A] -= 3;
  int mb[3] = {N[0], N[1], N[2]};
  mb[B] -= 3;
  auto at = [&](const int* dims, int cc, int aa, int bb) -> size_t {
    int p[3];
    p[CC] = cc;
    p[A] = aa;
    p[B] = bb;
    return ((size_t)p[0] * dims[1] + p[1]) * dims[2] + p[2];
  };
  const size_t FA = (size_t)ma[0] * ma[1] * ma[2];
  const size_t FB = (size_t)mb[0] * mb[1] * mb[2];
  const size_t P = g.cells();
  auto ecell = [&](int aa, int bb) -> double {
    const size_t c = at(N, kc, aa, bb);
    return prim[(5 + A) * P + c] * prim[(1 + B) * P + c] - prim[(5 + B) * P + c] * prim[(1 + A) * P + c];
  };
  const double v = ct_node(fa[5 * FA + at(ma, kc, i - 2, j - 1)], fa[5 * FA + at(ma, kc, i - 2, j)],
                           fb[6 * FB + at(mb, kc, i - 1, j - 2)], fb[6 * FB + at(mb, kc, i, j - 2)],
                           fa[at(ma, kc, i - 2, j - 1)], fa[at(ma, kc, i - 2, j)],
                           fb[at(mb, kc, i - 1, j - 2)], fb[at(mb, kc, i, j - 2)],
                           ecell(i - 1, j - 1), ecell(i, j - 1), ecell(i - 1, j), ecell(i, j));
  e[((size_t)o[0] * m[1] + o[1]) * m[2] + o[2]] = v;
}

// The stage update of the 5 cell fields, u' = u - (dt/dx dFx + dt/dy dFy + dt/dz dFz),
// and of the edge potential, A' = A - dt E, from `in` into `out` (with heun:
// out = (u' + out) * 0.5).
__global__ void k_update3d(const double* __restrict__ in, double* __restrict__ out,
                           const double* __restrict__ fx, const double* __restrict__ fy,
                           const double* __restrict__ fz, const double* __restrict__ ex,
                           const double* __restrict__ ey, const double* __restrict__ ez, Dims3 g,
                           double dx, double dy, double dz, const double* __restrict__ dtp,
                           int heun) {
  const double dt = *dtp;
  const double dtdx = dt / dx, dtdy = dt / dy, dtdz = dt / dz;
  const int nx = g.X - 4, ny = g.Y - 4, nz = g.Z - 4, h = HALO;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  const int c = t % (nz + 1), b = t / (nz + 1);
  if (a > nx || b > ny) return;
  const int X = g.X, Y = g.Y, Z = g.Z;
  const size_t P = g.cells();
  if (a < nx && b < ny && c < nz) {
    const size_t cell = ((size_t)(a + h) * Y + b + h) * Z + c + h;
    const size_t SX = (size_t)(X - 3) * Y * Z, SY = (size_t)X * (Y - 3) * Z, SZ = (size_t)X * Y * (Z - 3);
    const size_t ix0 = ((size_t)a * Y + b + h) * Z + c + h, ix1 = ix0 + (size_t)Y * Z;
    const size_t iy0 = ((size_t)(a + h) * (Y - 3) + b) * Z + c + h, iy1 = iy0 + Z;
    const size_t iz0 = ((size_t)(a + h) * Y + b + h) * (Z - 3) + c, iz1 = iz0 + 1;
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      int rx, ry, rz;
      if (f == 0) { rx = ry = rz = 0; }
      else if (f == 4) { rx = ry = rz = 4; }
      else {
        const int mm = f - 1;
        rx = 1 + (mm - 0 + 3) % 3;
        ry = 1 + (mm - 1 + 3) % 3;
        rz = 1 + (mm - 2 + 3) % 3;
      }
      const double dfx = fx[rx * SX + ix1] - fx[rx * SX + ix0];
      const double dfy = fy[ry * SY + iy1] - fy[ry * SY + iy0];
      const double dfz = fz[rz * SZ + iz1] - fz[rz * SZ + iz0];
      double v = in[f * P + cell] - (dtdx * dfx + dtdy * dfy + dtdz * dfz);
      if (heun) v = (v + out[f * P + cell]) * 0.5;
      out[f * P + cell] = v;
    }
  }
  // edges: A_x (X, Y+1, Z+1) interior (nx, ny+1, nz+1); E_x same interior shape
  const double* E[3] = {ex, ey, ez};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int lim[3] = {d == 0 ? nx : nx + 1, d == 1 ? ny : ny + 1, d == 2 ? nz : nz + 1};
    if (a >= lim[0] || b >= lim[1] || c >= lim[2]) continue;
    const int ext1 = d == 1 ? Y : Y + 1, ext2 = d == 2 ? Z : Z + 1;
    const size_t off = array_offset(g, 5 + d);
    const size_t idx = off + ((size_t)(a + h) * ext1 + b + h) * ext2 + c + h;
    double v = in[idx] - dt * E[d][((size_t)a * lim[1] + b) * lim[2] + c];
    if (heun) v = (v + out[idx]) * 0.5;
    out[idx] = v;
  }
}

// mhd3d.py:674-724: the largest sum over axes of (|u_a| + c_f,a) / dx_a and the
// smallest density of the block's own cells.
// One halo pass along `axis` over the 8 arrays of the state.  kind 0 periodic,
// 1 outflow (cells zero-gradient, potentials linearly extrapolated,
// mhd3d.py:448-473 with mhd.py:298-325).
__global__ void k_halo3d(double* __restrict__ state, Dims3 g, int axis, int kind) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int slot = blockIdx.y;
  const int arr = blockIdx.z;
  const int e0 = extent(g, arr, 0), e1 = extent(g, arr, 1), e2 = extent(g, arr, 2);
  const int e[3] = {e0, e1, e2};
  const int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;
  if (t >= e[o1] * e[o2]) return;
  const int q1 = t / e[o2], q2 = t % e[o2];
  const size_t st[3] = {(size_t)e1 * e2, (size_t)e2, 1};
  double* line = state + array_offset(g, arr) + q1 * st[o1] + q2 * st[o2];
  const size_t sa = st[axis];
  const int L = e[axis];
  const int n = (axis == 0 ? g.X : (axis == 1 ? g.Y : g.Z)) - 4;
  const bool node = arr >= 5 && arr - 5 != axis;
  const int h = HALO;
  auto at = [&](int idx) -> double& { return line[(size_t)idx * sa]; };
  if (kind == 0) {
    if (slot < 2) {
      at(slot) = at(slot + n);
    } else {
      const int tt = slot - 2;
      if (!node && tt >= 2) return;
      at(n + h + tt) = at(h + tt);
    }
  } else {
    if (slot >= 4) return;
    if (arr < 5) {
      if (slot < 2) at(slot) = at(h);
      else at(L - 4 + slot) = at(L - h - 1);
    } else {
      if (slot < 2) {
        const int k = h - slot;
        const double edge = at(h), inner = at(h + 1);
        at(h - k) = edge + (double)k * (edge - inner);
      } else {
        const int k = slot - 1;
        const double edge = at(L - h - 1), inner = at(L - h - 2);
        at(L - h - 1 + k) = edge + (double)k * (edge - inner);
      }
    }
  }
}

inline bool dims_ok(int64_t nx, int64_t ny, int64_t nz) {
  return nx >= 2 && ny >= 2 && nz >= 2 && nx < 65536 && ny < 65536 && nz < 65536 &&
         (nx + 5) * (ny + 5) * (nz + 5) < (int64_t)1 << 36;
}

}  // namespace

extern "C" {

int64_t wb_mhd3d_state_size(int64_t nx, int64_t ny, int64_t nz) {
  if (!dims_ok(nx, ny, nz)) return WB_EINVAL;
  Dims3 g{(int)nx + 4, (int)ny + 4, (int)nz + 4};
  return (int64_t)array_offset(g, 8);
}

}  // extern "C"

namespace {

// The halo passes of the axes in `axes`: periodic axes first, then outflow axes.
int fill_halos3d(double* state, int64_t nx, int64_t ny, int64_t nz, int periodic_mask, int axes,
                 hipStream_t s) {
  if (!dims_ok(nx, ny, nz)) return WB_EINVAL;
  Dims3 g{(int)nx + 4, (int)ny + 4, (int)nz + 4};
  const int64_t n[3] = {g.X + 1, g.Y + 1, g.Z + 1};
  for (int pass = 0; pass < 2; ++pass)
    for (int axis = 0; axis < 3; ++axis) {
      const bool periodic = (periodic_mask >> axis) & 1;
      if ((pass == 0) != periodic || !((axes >> axis) & 1)) continue;
      const int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;
      const int64_t other = n[o1] * n[o2];
      dim3 grid((unsigned)((other + 255) / 256), 5, 8);
      hipLaunchKernelGGL(k_halo3d, grid, dim3(256), 0, s, state, g, axis, periodic ? 0 : 1);
    }
  return wb_last_error();
}

}  // namespace

extern "C" {

int wb_mhd3d_fill_halos(double* state, int64_t nx, int64_t ny, int64_t nz, int periodic_mask,
                        void* stream) {
  return fill_halos3d(state, nx, ny, nz, periodic_mask, 7, (hipStream_t)stream);
}

int wb_mhd3d_fill_outflow(double* state, int64_t nx, int64_t ny, int64_t nz, int periodic_mask,
                          void* stream) {
  return fill_halos3d(state, nx, ny, nz, periodic_mask, ~periodic_mask & 7, (hipStream_t)stream);
}

int wb_primitive3d(const double* state, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                   double dz, double b0x, double b0y, double b0z, double gamma, double* prim,
                   double* fbx, double* fby, double* fbz, int32_t* bad, void* stream) {
  if (!dims_ok(nx, ny, nz)) return WB_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  Dims3 g{(int)nx + 4, (int)ny + 4, (int)nz + 4};
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, bad, 1);
  const int64_t inner = (int64_t)(g.Y + 1) * (g.Z + 1);
  dim3 grid((unsigned)((inner + 255) / 256), (unsigned)(g.X + 1));
  hipLaunchKernelGGL(k_primitive3d, grid, dim3(256), 0, s, state, g, dx, dy, dz, b0x, b0y, b0z,
                     gamma, prim, fbx, fby, fbz, (int*)bad);
  return wb_last_error();
}

int wb_sweep3d(const double* prim, const double* bn, int64_t nx, int64_t ny, int64_t nz, int axis,
               int limiter, double gamma, double* flux, void* stream) {
  if (!dims_ok(nx, ny, nz) || axis < 0 || axis > 2 || limiter < 0 || limiter > 3) return WB_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  Dims3 g{(int)nx + 4, (int)ny + 4, (int)nz + 4};
  int m[3] = {g.X, g.Y, g.Z};
  m[axis] -= 3;
  const int64_t inner = (int64_t)m[1] * m[2];
  dim3 grid((unsigned)((inner + 127) / 128), (unsigned)m[0]);
  if (axis == 0)
    hipLaunchKernelGGL(k_sweep3d<0>, grid, dim3(128), 0, s, prim, bn, g, limiter, gamma, flux);
  else if (axis == 1)
    hipLaunchKernelGGL(k_sweep3d<1>, grid, dim3(128), 0, s, prim, bn, g, limiter, gamma, flux);
  else
    hipLaunchKernelGGL(k_sweep3d<2>, grid, dim3(128), 0, s, prim, bn, g, limiter, gamma, flux);
  return wb_last_error();
}

int wb_edge3d(const double* prim, const double* flux_a, const double* flux_b, int64_t nx,
              int64_t ny, int64_t nz, int c, double* e, void* stream) {
  if (!dims_ok(nx, ny, nz) || c < 0 || c > 2) return WB_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  Dims3 g{(int)nx + 4, (int)ny + 4, (int)nz + 4};
  int m[3] = {(int)nx + 1, (int)ny + 1, (int)nz + 1};
  m[c] -= 1;
  const int64_t inner = (int64_t)m[1] * m[2];
  dim3 grid((unsigned)((inner + 255) / 256), (unsigned)m[0]);
  if (c == 0) hipLaunchKernelGGL(k_edge3d<0>, grid, dim3(256), 0, s, prim, flux_a, flux_b, g, e);
  else if (c == 1) hipLaunchKernelGGL(k_edge3d<1>, grid, dim3(256), 0, s, prim, flux_a, flux_b, g, e);
  else hipLaunchKernelGGL(k_edge3d<2>, grid, dim3(256), 0, s, prim, flux_a, flux_b, g, e);
  return wb_last_error();
}

int wb_mhd3d_update(const double* in, double* out, const double* fx, const double* fy,
                    const double* fz, const double* ex, const double* ey, const double* ez,
                    int64_t nx, int64_t ny, int64_t nz, const double* dt, double dx, double dy,
                    double dz, int heun, void* stream) {
  if (!dims_ok(nx, ny, nz)) return WB_EINVAL;
  Dims3 g{(int)nx + 4, (int)ny + 4, (int)nz + 4};
  const int64_t inner = (ny + 1) * (nz + 1);
  dim3 grid((unsigned)((inner + 255) / 256), (unsigned)(nx + 1));
  hipLaunchKernelGGL(k_update3d, grid, dim3(256), 0, (hipStream_t)stream, in, out, fx, fy, fz, ex,
                     ey, ez, g, dx, dy, dz, dt, heun);
  return wb_last_error();
}

int wb_rate3d(const double* state, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
              double dz, double b0x, double b0y, double b0z, double gamma, double* out,
              void* stream) {
  if (!dims_ok(nx, ny, nz)) return WB_EINVAL;
  return launch_rate3d<false>(state, nx, ny, nz, dx, dy, dz, b0x, b0y, b0z, gamma, out,
                              (hipStream_t)stream);
}

}  // extern "C"
