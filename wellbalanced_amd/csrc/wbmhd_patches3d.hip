// The refined 3D step on patch hierarchies of boxes of any size (MHD3D on patchwork
// patches, carafe mhd3d.py:448-473,561-602,635-671,726-754): the table-driven ghost
// fill, average-down and restriction (one gather kernel), the flux register and reflux
// from the split path's sweep arrays, the resistive terms added into a patch's sweeps
// and edge fields before its update, and the moves of a regrid.  FP64,
// -ffp-contract=off: every operation rounds as NumPy's does.
#include "wbmhd_common.hpp"

namespace {

inline unsigned grid_of(long long n) { return (unsigned)((n + 255) / 256); }

// dst[i] = w0 src[s0] + w1 src[s1] + ... in term order (the gather tables of
// wellbalanced_amd/patches3d_maps.py: ghosts, average-down, restriction, regrid moves).
// No destination is a source of the same launch.
__global__ void k_gather(double* __restrict__ dst_buf, const double* __restrict__ src_buf,
                         const long long* __restrict__ dst, const long long* __restrict__ start,
                         const int* __restrict__ count, const long long* __restrict__ src,
                         const double* __restrict__ w, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long t0 = start[i];
  const int m = count[i];
  double v = w[t0] * src_buf[src[t0]];
  for (int t = 1; t < m; ++t) v = v + w[t0 + t] * src_buf[src[t0 + t]];
  dst_buf[dst[i]] = v;
}

__device__ __forceinline__ double minmod(double a, double b) {
  if (!(a * b > 0.0)) return 0.0;
  return __builtin_fabs(a) < __builtin_fabs(b) ? a : b;
}

// A new cell of a regrid: the coarse cell's value plus its minmod slopes towards the
// child, c0 + sx/4 slope_x + sy/4 slope_y + sz/4 slope_z in that order, per field
// (the children of a coarse cell sum to 8 times it).  rows: dst, dst plane, coarse
// cell, coarse x stride, coarse y stride, coarse plane, signs (sx, sy, sz as +-1).
__global__ void k_prolong_cells3(double* __restrict__ buf, const long long* __restrict__ rows,
                                 long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long* r = rows + 9 * i;
  const long long d = r[0], pd = r[1], c = r[2], sx = r[3], sy = r[4], pc = r[5];
  const double hx = 0.25 * (double)r[6], hy = 0.25 * (double)r[7], hz = 0.25 * (double)r[8];
  for (int f = 0; f < 5; ++f) {
    const long long q = c + f * pc;
    const double c0 = buf[q];
    const double slx = minmod(buf[q + sx] - c0, c0 - buf[q - sx]);
    const double sly = minmod(buf[q + sy] - c0, c0 - buf[q - sy]);
    const double slz = minmod(buf[q + 1] - c0, c0 - buf[q - 1]);
    buf[d + f * pd] = ((c0 + hx * slx) + hy * sly) + hz * slz;
  }
}

// The flux register of the coarse-fine faces (FluxRegister.update with scale 0.5 dt,
// FluxRegister.reflux; mhd3d.py:650-655,745-750), from the sweep arrays of the stage:
// per refluxed coarse cell g, d = sum over its faces (x-high, x-low, ...) of
// dt/h (F_c - mean) (high face) or dt/h (mean - F_c) (low face), mean =
// ((f00 + f01) + (f10 + f11)) / 4.  record: reg[5 g + f] = 0.5 d; else the cell gains
// reg + 0.5 d.  cells: {cell offset, cell plane, first face, face count} (int64);
// faces: {axis, side, coarse flux offset, row stride, then the four fine ones} (int64)
// and the coarse spacing along the axis (double).
__global__ void k_reflux_patches3(double* __restrict__ out, const double* __restrict__ sweeps,
                                  const long long* __restrict__ cells,
                                  const long long* __restrict__ faces,
                                  const double* __restrict__ spacing, long long ncells,
                                  const double* __restrict__ dtp, int record,
                                  double* __restrict__ reg) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncells) return;
  constexpr int rows[3][5] = {{0, 1, 2, 3, 4}, {0, 3, 1, 2, 4}, {0, 2, 3, 1, 4}};
  const long long* c = cells + 4 * g;
  const double dt = *dtp;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (long long k = c[2]; k < c[2] + c[3]; ++k) {
    const long long* f = faces + 12 * k;
    const int axis = (int)f[0], side = (int)f[1];
    const double dtd = dt / spacing[k];
    for (int fld = 0; fld < 5; ++fld) {
      const int r = rows[axis][fld];
      const double fc = sweeps[f[2] + r * f[3]];
      const double f00 = sweeps[f[4] + r * f[5]], f01 = sweeps[f[6] + r * f[7]];
      const double f10 = sweeps[f[8] + r * f[9]], f11 = sweeps[f[10] + r * f[11]];
      const double fm = ((f00 + f01) + (f10 + f11)) * 0.25;
      const double d = side == 0 ? dtd * (fc - fm) : dtd * (fm - fc);
      acc[fld] = acc[fld] + d;
    }
  }
  double* kept = reg + 5 * g;
  if (record) {
    for (int fld = 0; fld < 5; ++fld) kept[fld] = 0.5 * acc[fld];
    return;
  }
  for (int fld = 0; fld < 5; ++fld) {
    double* u = out + c[0] + fld * c[1];
    *u = *u + (kept[fld] + 0.5 * acc[fld]);
  }
}

// MHD3D._resistive (mhd3d.py:561-602) in the split path, before the update: J_c on the
// block's c-edges from the haloed face fields, J_c = d_a B_b / h_a + (-(d_b B_a / h_b)),
// E_c += eta J_c; and on a face normal to s the energy flux (row 4 of the sweep along s)
// gains eta (J_b B_c - J_c B_b), J the mean of the face's two edges along each direction
// and B the mean of the face's four faces around it (normal direction first).
struct Block3 {
  int nx, ny, nz, X, Y, Z;
  __device__ int n(int a) const { return a == 0 ? nx : (a == 1 ? ny : nz); }
  __device__ int H(int a) const { return a == 0 ? X : (a == 1 ? Y : Z); }
};

// Haloed face field of axis a (node-centred along a) at haloed indices p[3].
__device__ __forceinline__ double face_at(const double* f, const Block3& g, int a, const int* p) {
  const long long ey = g.Y + (a == 1), ez = g.Z + (a == 2);
  return f[((long long)p[0] * ey + p[1]) * ez + p[2]];
}

// J_c at block c-edge t[3] (along c: cell 0..n_c-1; along the others: node 0..n).
__device__ double current_at(const double* const* F, const Block3& g, int c, const int* t,
                             double h0, double h1, double h2) {
  const int a = (c + 1) % 3, b = (c + 2) % 3;
  const double hs[3] = {h0, h1, h2};
  int p[3] = {2 + t[0], 2 + t[1], 2 + t[2]};
  int m[3] = {p[0], p[1], p[2]};
  m[a] -= 1;
  const double part0 = (face_at(F[b], g, b, p) - face_at(F[b], g, b, m)) / hs[a];
  int m2[3] = {p[0], p[1], p[2]};
  m2[b] -= 1;
  const double part1 = -((face_at(F[a], g, a, p) - face_at(F[a], g, a, m2)) / hs[b]);
  return part0 + part1;
}

__global__ void k_resistive_edges3(const double* __restrict__ f0, const double* __restrict__ f1,
                                   const double* __restrict__ f2, Block3 g, double h0, double h1,
                                   double h2, double eta, double* __restrict__ e0,
                                   double* __restrict__ e1, double* __restrict__ e2) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  const double* F[3] = {f0, f1, f2};
  double* E[3] = {e0, e1, e2};
  int ext[3] = {g.nx + 1, g.ny + 1, g.nz + 1};
  ext[c] -= 1;
  if (t >= (long long)ext[0] * ext[1] * ext[2]) return;
  int p[3];
  p[2] = (int)(t % ext[2]);
  p[1] = (int)((t / ext[2]) % ext[1]);
  p[0] = (int)(t / ((long long)ext[2] * ext[1]));
  const double j = current_at(F, g, c, p, h0, h1, h2);
  double* e = E[c] + t;
  *e = *e + eta * j;
}

// The face-centred means of B_c on the s-faces (mhd3d.py:105-117): along s the mean of
// the two cells' faces beside the face, then along the node axis c the mean of the two
// node positions around the cell centre.
__device__ double field_at_face(const double* F, const Block3& g, int comp, int s, int node,
                                const int* q) {
  auto at = [&](int ds, int dn) {
    int p[3] = {2 + q[0], 2 + q[1], 2 + q[2]};
    p[s] += ds - 1;  // the face at block position q[s] lies between cells q[s]-1 and q[s]
    p[node] += dn;
    return face_at(F, g, comp, p);
  };
  const double lo = 0.5 * (at(0, 0) + at(1, 0));
  const double hi = 0.5 * (at(0, 1) + at(1, 1));
  return 0.5 * (lo + hi);
}

__global__ void k_resistive_energy3(const double* __restrict__ f0, const double* __restrict__ f1,
                                    const double* __restrict__ f2, Block3 g, double h0, double h1,
                                    double h2, double eta, double* __restrict__ s0,
                                    double* __restrict__ s1, double* __restrict__ s2) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  const double* F[3] = {f0, f1, f2};
  double* S[3] = {s0, s1, s2};
  int ext[3] = {g.nx, g.ny, g.nz};
  ext[s] += 1;
  if (t >= (long long)ext[0] * ext[1] * ext[2]) return;
  int q[3];
  q[2] = (int)(t % ext[2]);
  q[1] = (int)((t / ext[2]) % ext[1]);
  q[0] = (int)(t / ((long long)ext[2] * ext[1]));
  const int b = (s + 1) % 3, c = (s + 2) % 3;
  // J_b on the face: the mean of the b-edges at nodes q[c], q[c] + 1 (q[s] a node)
  int e[3] = {q[0], q[1], q[2]};
  const double jb0 = current_at(F, g, b, e, h0, h1, h2);
  e[c] += 1;
  const double jb1 = current_at(F, g, b, e, h0, h1, h2);
  const double jb = 0.5 * (jb0 + jb1);
  int e2[3] = {q[0], q[1], q[2]};
  const double jc0 = current_at(F, g, c, e2, h0, h1, h2);
  e2[b] += 1;
  const double jc1 = current_at(F, g, c, e2, h0, h1, h2);
  const double jc = 0.5 * (jc0 + jc1);
  const double bc = field_at_face(F[c], g, c, s, c, q);
  const double bb = field_at_face(F[b], g, b, s, b, q);
  // the sweep along s: (7, n_s + 1 faces, haloed others), face index q[s]
  long long dims[3] = {g.X, g.Y, g.Z};
  dims[s] = g.n(s) + 1;
  int idx[3] = {2 + q[0], 2 + q[1], 2 + q[2]};
  idx[s] = q[s];
  const long long plane = dims[0] * dims[1] * dims[2];
  double* en = S[s] + 4 * plane + ((long long)idx[0] * dims[1] + idx[1]) * dims[2] + idx[2];
  *en = *en + eta * (jb * bc - jc * bb);
}

}  // namespace

extern "C" {

int wb_patches3d_gather(double* dst_buf, const double* src_buf, const long long* dst,
                        const long long* start, const int32_t* count, const long long* src,
                        const double* w, int64_t n, void* stream) {
  if (n < 0) return WB_EINVAL;
  if (n == 0) return 0;
  if (!dst_buf || !src_buf || !dst || !start || !count || !src || !w) return WB_EINVAL;
  hipLaunchKernelGGL(k_gather, dim3(grid_of(n)), dim3(256), 0, (hipStream_t)stream, dst_buf,
                     src_buf, dst, start, (const int*)count, src, w, (long long)n);
  return wb_last_error();
}

int wb_patches3d_prolong(double* buf, const long long* rows, int64_t n, void* stream) {
  if (n < 0) return WB_EINVAL;
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_prolong_cells3, dim3(grid_of(n)), dim3(256), 0, (hipStream_t)stream, buf,
                     rows, (long long)n);
  return wb_last_error();
}

int wb_patches3d_reflux(double* out, const double* sweeps, const long long* cells,
                        const long long* faces, const double* spacing, int64_t ncells,
                        const double* dt, int record, double* reg, void* stream) {
  if (ncells < 0) return WB_EINVAL;
  if (ncells == 0) return 0;
  if (!reg || !dt || !cells || !faces || !spacing || !sweeps) return WB_EINVAL;
  hipLaunchKernelGGL(k_reflux_patches3, dim3(grid_of(ncells)), dim3(256), 0, (hipStream_t)stream,
                     out, sweeps, cells, faces, spacing, (long long)ncells, dt, record, reg);
  return wb_last_error();
}

int wb_resistive3d(const double* f0, const double* f1, const double* f2, int64_t nx, int64_t ny,
                   int64_t nz, double dx, double dy, double dz, double eta, double* s0,
                   double* s1, double* s2, double* e0, double* e1, double* e2, void* stream) {
  if (nx < 1 || ny < 1 || nz < 1 || nx > 4096 || ny > 4096 || nz > 4096 || !(eta >= 0.0))
    return WB_EINVAL;
  if (eta == 0.0) return 0;
  const Block3 g{(int)nx, (int)ny, (int)nz, (int)nx + 4, (int)ny + 4, (int)nz + 4};
  hipStream_t s = (hipStream_t)stream;
  const long long edges = (nx + 1) * (ny + 1) * (nz + 1);
  const long long faces = (nx + 1) * (ny + 1) * (nz + 1);
  // the energy fluxes read J from the face fields, not from the edges: the two passes
  // are independent
  hipLaunchKernelGGL(k_resistive_energy3, dim3(grid_of(faces), 3), dim3(256), 0, s, f0, f1, f2, g,
                     dx, dy, dz, eta, s0, s1, s2);
  hipLaunchKernelGGL(k_resistive_edges3, dim3(grid_of(edges), 3), dim3(256), 0, s, f0, f1, f2, g,
                     dx, dy, dz, eta, e0, e1, e2);
  return wb_last_error();
}

}  // extern "C"
