// Gather kernels of the 2D patch hierarchy: halo fill from same-level patches and
// from the level below, the average of covered coarse cells and the restriction of
// A_z.  Each consumes a table of int64 offsets into the one FP64 buffer that holds
// every patch (built on the host from the hierarchy's geometry).  FP64,
// -ffp-contract=off.
#include "mhd_device.hpp"
#include "wbmhd_common.hpp"

using namespace wb;

namespace {

__global__ void k_copy_cells(double* __restrict__ buf, const long long* __restrict__ dst,
                             const long long* __restrict__ src, const long long* __restrict__ pd,
                             const long long* __restrict__ ps, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long d = dst[i], s = src[i], a = pd[i], b = ps[i];
#pragma unroll
  for (int f = 0; f < 6; ++f) buf[d + f * a] = buf[s + f * b];
}

__global__ void k_copy_nodes(double* __restrict__ buf, const long long* __restrict__ dst,
                             const long long* __restrict__ src, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) buf[dst[i]] = buf[src[i]];
}

// Bilinear from the 2x2 nearest coarse cells; wx1, wy1: weights of the upper cells.
__global__ void k_interp_cells(double* __restrict__ buf, const long long* __restrict__ dst,
                               const long long* __restrict__ pd, const long long* __restrict__ c00,
                               const long long* __restrict__ cy, const long long* __restrict__ pc,
                               const double* __restrict__ wx, const double* __restrict__ wy, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double wx1 = wx[i], wy1 = wy[i];
  const double wx0 = 1.0 - wx1, wy0 = 1.0 - wy1;
  const long long d = dst[i], a = pd[i], sy = cy[i], p = pc[i];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const long long b = c00[i] + f * p;
    buf[d + f * a] = wy0 * (wx0 * buf[b] + wx1 * buf[b + sy]) + wy1 * (wx0 * buf[b + 1] + wx1 * buf[b + sy + 1]);
  }
}

// Per axis: the coarse node itself (kind 0) or the 4-point midpoint rule
// (-1, 9, 9, -1) / 16 on the nodes base .. base+3 (kind 1); a tensor product.
__global__ void k_interp_nodes(double* __restrict__ buf, const long long* __restrict__ dst,
                               const long long* __restrict__ base, const long long* __restrict__ sx,
                               const long long* __restrict__ sy, const long long* __restrict__ kx,
                               const long long* __restrict__ ky, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double w4[4] = {-0.0625, 0.5625, 0.5625, -0.0625};
  const int nxw = kx[i] ? 4 : 1, nyw = ky[i] ? 4 : 1;
  double total = 0.0;
  for (int j = 0; j < nyw; ++j) {
    double row = 0.0;
    for (int k = 0; k < nxw; ++k) {
      const double wk = kx[i] ? w4[k] : 1.0;
      row = row + wk * buf[base[i] + k * sx[i] + j * sy[i]];
    }
    const double wj = ky[i] ? w4[j] : 1.0;
    total = total + wj * row;
  }
  buf[dst[i]] = total;
}

// Regridding: the cells (6 fields) and nodes a patch of the same level held in the old
// hierarchy's buffer, copied into the new buffer.
__global__ void k_regrid_copy_cells(const double* __restrict__ old, double* __restrict__ buf,
                                    const long long* __restrict__ dst,
                                    const long long* __restrict__ src,
                                    const long long* __restrict__ pd,
                                    const long long* __restrict__ ps, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long d = dst[i], s = src[i], a = pd[i], b = ps[i];
#pragma unroll
  for (int f = 0; f < 6; ++f) buf[d + f * a] = old[s + f * b];
}

__global__ void k_regrid_copy_nodes(const double* __restrict__ old, double* __restrict__ buf,
                                    const long long* __restrict__ dst,
                                    const long long* __restrict__ src, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) buf[dst[i]] = old[src[i]];
}

__device__ __forceinline__ double minmod2(double a, double b) {
  if (!(a * b > 0.0)) return 0.0;
  return __builtin_fabs(a) < __builtin_fabs(b) ? a : b;
}

// Conservative linear prolongation of a new fine cell from its coarse cell c (row
// length cy, plane pc): c + 0.25 sx s_x + 0.25 sy s_y with minmod slopes, sx/sy = -1
// for the lower child, +1 for the upper one; the four children sum to 4 c.
__global__ void k_prolong_cells(double* __restrict__ buf, const long long* __restrict__ dst,
                                const long long* __restrict__ pd, const long long* __restrict__ c,
                                const long long* __restrict__ cy, const long long* __restrict__ pc,
                                const long long* __restrict__ sx, const long long* __restrict__ sy,
                                int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long d = dst[i], a = pd[i], row = cy[i], p = pc[i];
  const double hx = 0.25 * (double)sx[i], hy = 0.25 * (double)sy[i];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const double* q = buf + c[i] + f * p;
    const double c0 = q[0];
    const double slx = minmod2(q[row] - c0, c0 - q[-row]);
    const double sly = minmod2(q[1] - c0, c0 - q[-1]);
    buf[d + f * a] = (c0 + hx * slx) + hy * sly;
  }
}

// Bilinear prolongation of a new fine node from the coarse nodes base, base + sx
// (kind x 1), base + 1 (kind y 1): the coarse node, the mean of two or of four.
__global__ void k_prolong_nodes(double* __restrict__ buf, const long long* __restrict__ dst,
                                const long long* __restrict__ base, const long long* __restrict__ sx,
                                const long long* __restrict__ kx, const long long* __restrict__ ky,
                                int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* q = buf + base[i];
  const long long s = sx[i];
  double v;
  if (kx[i] && ky[i])
    v = 0.25 * ((q[0] + q[s]) + (q[1] + q[s + 1]));
  else if (kx[i])
    v = 0.5 * (q[0] + q[s]);
  else if (ky[i])
    v = 0.5 * (q[0] + q[1]);
  else
    v = q[0];
  buf[dst[i]] = v;
}

// Covered coarse cell = mean of its four fine children, ((f00 + f10) + (f01 + f11)) / 4.
__global__ void k_average(double* __restrict__ buf, const long long* __restrict__ rows, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long* r = rows + 5 * (size_t)i;
  const long long coarse = r[0], pc = r[1], fine = r[2], fy = r[3], pf = r[4];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const long long b = fine + f * pf;
    buf[coarse + f * pc] = ((buf[b] + buf[b + fy]) + (buf[b + 1] + buf[b + fy + 1])) * 0.25;
  }
}

// A coarse node that coincides with a fine node takes the fine value (mhd.py:981,991).
__global__ void k_restrict(double* __restrict__ buf, const long long* __restrict__ rows, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) buf[rows[2 * (size_t)i]] = buf[rows[2 * (size_t)i + 1]];
}

// The CFL rate over the own cells of every patch (mhd_rate_pyccel.py:30-83 per patch,
// the largest over all).  The cells are cut into chunks of 256 of one patch (chunks:
// int32 pairs patch, first cell, built on the host); a fixed grid walks the chunks
// and each workgroup ends with one atomic pair.
__global__ void __launch_bounds__(256) k_rate_amr(const double* __restrict__ buf,
                                                  const PatchDesc* __restrict__ patches,
                                                  const int* __restrict__ chunks, int nchunks,
                                                  double b0x, double b0y, double gamma, double eta,
                                                  double cond, const double* __restrict__ spatch,
                                                  unsigned long long* keys) {
  double rate = 0.0, r = __builtin_inf(), ds = 0.0;
  for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const PatchDesc pd = patches[chunks[2 * ch]];
    const int t = chunks[2 * ch + 1] + (int)threadIdx.x;
    if (t >= pd.nx * pd.ny) continue;
    const int X = pd.nx + 4, Y = pd.ny + 4, YN = Y + 1;
    const int i = t / pd.ny + 2, j = t % pd.ny + 2;
    const double* in = buf + pd.off;
    const double* az = in + 6 * (size_t)X * Y;
    const double dx = pd.dx, dy = pd.dy;
    const size_t c = (size_t)i * Y + j, P = (size_t)X * Y;
    const double rc = in[c];
    const double u = in[P + c] / rc, v = in[2 * P + c] / rc, w = in[3 * P + c] / rc;
    const double bx = 0.5 * ((az[(size_t)i * YN + j + 1] - az[(size_t)i * YN + j]) / dy + b0x +
                             ((az[(size_t)(i + 1) * YN + j + 1] - az[(size_t)(i + 1) * YN + j]) / dy + b0x));
    const double by = 0.5 * (-(az[(size_t)(i + 1) * YN + j] - az[(size_t)i * YN + j]) / dx + b0y +
                             (-(az[(size_t)(i + 1) * YN + j + 1] - az[(size_t)i * YN + j + 1]) / dx + b0y));
    const double bzc = in[5 * P + c];
    const double p = (gamma - 1.0) * (in[4 * P + c] - 0.5 * rc * (u * u + v * v + w * w) -
                                      0.5 * (bx * bx + by * by + bzc * bzc));
    const double cr = (__builtin_fabs(u) + fast_speed_rate(rc, p, bx, by, bzc, gamma)) / dx +
                      (__builtin_fabs(v) + fast_speed_rate(rc, p, by, bx, bzc, gamma)) / dy;
    // NaN rates and densities drop out, as in the single-patch reduction
    rate = cr > rate ? cr : rate;
    r = rc < r ? rc : r;
    if (spatch) {  // the diffusion number D s of the cell's patch (mhd.py:955-961)
      const double c = cond / rc;
      const double d = (c > eta ? c : eta) * spatch[chunks[2 * ch]];
      ds = d > ds ? d : ds;
    }
  }
  if (spatch) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(ds, off, 64);
      ds = o > ds ? o : ds;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(&keys[2], order_key(ds));
  }
  block_reduce_rate_rho_once(rate, r, keys);
}

// keys[2] (the largest D s) starts at key(0); finished into out[2].
__global__ void k_rate3_init(unsigned long long* keys) {
  if (threadIdx.x == 0) keys[2] = order_key(0.0);
}
__global__ void k_rate3_finish(double* out) {
  if (threadIdx.x == 0) {
    const unsigned long long k = ((const unsigned long long*)out)[2];
    out[2] = key_to_double(k);
  }
}

inline dim3 blocks(int n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

extern "C" {

int wb_amr2d_copy(double* buf, const long long* cdst, const long long* csrc, const long long* cpd,
                  const long long* cps, int ncells, const long long* ndst, const long long* nsrc,
                  int nnodes, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (ncells > 0) hipLaunchKernelGGL(k_copy_cells, blocks(ncells), dim3(256), 0, s, buf, cdst, csrc, cpd, cps, ncells);
  if (nnodes > 0) hipLaunchKernelGGL(k_copy_nodes, blocks(nnodes), dim3(256), 0, s, buf, ndst, nsrc, nnodes);
  return wb_last_error();
}

int wb_amr2d_interp(double* buf, const long long* dst, const long long* pd, const long long* c00,
                    const long long* cy, const long long* pc, const double* wx, const double* wy,
                    int ncells, const long long* ndst, const long long* nbase,
                    const long long* nsx, const long long* nsy, const long long* nkx,
                    const long long* nky, int nnodes, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (ncells > 0) hipLaunchKernelGGL(k_interp_cells, blocks(ncells), dim3(256), 0, s, buf, dst, pd, c00, cy, pc, wx, wy, ncells);
  if (nnodes > 0) hipLaunchKernelGGL(k_interp_nodes, blocks(nnodes), dim3(256), 0, s, buf, ndst, nbase, nsx, nsy, nkx, nky, nnodes);
  return wb_last_error();
}

int wb_amr2d_regrid_copy(const double* old, double* buf, const long long* cdst,
                         const long long* csrc, const long long* cpd, const long long* cps,
                         int ncells, const long long* ndst, const long long* nsrc, int nnodes,
                         void* stream) {
  if (ncells < 0 || nnodes < 0) return WB_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  if (ncells > 0) hipLaunchKernelGGL(k_regrid_copy_cells, blocks(ncells), dim3(256), 0, s, old, buf, cdst, csrc, cpd, cps, ncells);
  if (nnodes > 0) hipLaunchKernelGGL(k_regrid_copy_nodes, blocks(nnodes), dim3(256), 0, s, old, buf, ndst, nsrc, nnodes);
  return wb_last_error();
}

int wb_amr2d_prolong(double* buf, const long long* dst, const long long* pd, const long long* c,
                     const long long* cy, const long long* pc, const long long* sx,
                     const long long* sy, int ncells, const long long* ndst,
                     const long long* nbase, const long long* nsx, const long long* nkx,
                     const long long* nky, int nnodes, void* stream) {
  if (ncells < 0 || nnodes < 0) return WB_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  if (ncells > 0) hipLaunchKernelGGL(k_prolong_cells, blocks(ncells), dim3(256), 0, s, buf, dst, pd, c, cy, pc, sx, sy, ncells);
  if (nnodes > 0) hipLaunchKernelGGL(k_prolong_nodes, blocks(nnodes), dim3(256), 0, s, buf, ndst, nbase, nsx, nkx, nky, nnodes);
  return wb_last_error();
}

int wb_amr2d_average(double* buf, const long long* rows, int n, void* stream) {
  if (n > 0) hipLaunchKernelGGL(k_average, blocks(n), dim3(256), 0, (hipStream_t)stream, buf, rows, n);
  return wb_last_error();
}

int wb_amr2d_restrict(double* buf, const long long* rows, int n, void* stream) {
  if (n > 0) hipLaunchKernelGGL(k_restrict, blocks(n), dim3(256), 0, (hipStream_t)stream, buf, rows, n);
  return wb_last_error();
}

int wb_amr2d_rate(const double* buf, const void* patches, const int32_t* chunks, int nchunks,
                  double b0x, double b0y, double gamma, double eta, double cond,
                  const double* spatch, double* out, void* stream) {
  if (nchunks < 1 || !(eta >= 0.0) || !(cond >= 0.0)) return WB_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  unsigned long long* keys = (unsigned long long*)out;
  hipLaunchKernelGGL(k_rate_init, dim3(1), dim3(64), 0, s, keys);
  if (spatch) hipLaunchKernelGGL(k_rate3_init, dim3(1), dim3(64), 0, s, keys);
  const unsigned grid = (unsigned)(nchunks < 1024 ? nchunks : 1024);
  hipLaunchKernelGGL(k_rate_amr, dim3(grid), dim3(256), 0, s, buf, (const PatchDesc*)patches,
                     (const int*)chunks, nchunks, b0x, b0y, gamma, eta, cond, spatch, keys);
  hipLaunchKernelGGL(k_rate_finish, dim3(1), dim3(64), 0, s, keys, out);
  if (spatch) hipLaunchKernelGGL(k_rate3_finish, dim3(1), dim3(64), 0, s, out);
  return wb_last_error();
}

}  // extern "C"
