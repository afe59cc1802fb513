// Gather kernels of the 2D patch hierarchy: halo fill from same-level patches and
// from the level below, the average of covered coarse cells and the restriction of
// A_z.  Each consumes a table of int64 offsets into the one FP64 buffer that holds
// every patch (built on the host from the hierarchy's geometry).  FP64,
// -ffp-contract=off.
#include "mhd_device.hpp"
#include "wbmhd_common.hpp"

using namespace wb;

namespace {

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
__host__ __device__ inline int nblocks(int n) { return n > 0 ? (n + 255) / 256 : 0; }

// The tables of one level's halo fill: same-level copies of cells and nodes, then the
// interpolation of cells and nodes from the level below.
struct FillLevel {
  const long long *cdst, *csrc, *cpd, *cps;
  const long long *ndst, *nsrc;
  const long long *idst, *ipd, *ic00, *icy, *ipc;
  const double *iwx, *iwy;
  const long long *jdst, *jbase, *jsx, *jsy, *jkx, *jky;
  int ncells, nnodes, nic, nin;
};

// One level's halo fill in one launch, the workgroups cut into four parts of 256
// threads each: ghost cells (6 fields) and ghost or shared boundary A_z nodes copied
// from a same-level owner; the other ghost cells bilinear from the 2x2 nearest coarse
// cells (wx1, wy1: weights of the upper cells); the other ghost nodes, per axis the
// coarse node itself (kind 0) or, midway between coarse nodes m and m+1, the quadratic
// through nodes m-1, m, m+1, weights (-1/8, 3/4, 3/8) on base .. base+2 (kind 1), a
// tensor product (the reference's quadratic potential halos, CHANGELOG.md:10-13).  The
// parts write disjoint ghosts and read only interior values of the level or values of
// the level below, so they need no order among themselves.
__global__ void __launch_bounds__(256) k_fill_level(double* __restrict__ buf, FillLevel t) {
  int b = blockIdx.x;
  const int b0 = nblocks(t.ncells), b1 = nblocks(t.nnodes), b2 = nblocks(t.nic);
  const int tid = threadIdx.x;
  if (b < b0) {
    const int i = b * 256 + tid;
    if (i >= t.ncells) return;
    const long long d = t.cdst[i], sr = t.csrc[i], a = t.cpd[i], q = t.cps[i];
#pragma unroll
    for (int f = 0; f < 6; ++f) buf[d + f * a] = buf[sr + f * q];
    return;
  }
  b -= b0;
  if (b < b1) {
    const int i = b * 256 + tid;
    if (i < t.nnodes) buf[t.ndst[i]] = buf[t.nsrc[i]];
    return;
  }
  b -= b1;
  if (b < b2) {
    const int i = b * 256 + tid;
    if (i >= t.nic) return;
    const double wx1 = t.iwx[i], wy1 = t.iwy[i];
    const double wx0 = 1.0 - wx1, wy0 = 1.0 - wy1;
    const long long d = t.idst[i], a = t.ipd[i], sy = t.icy[i], pc = t.ipc[i];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const long long q = t.ic00[i] + f * pc;
      buf[d + f * a] = wy0 * (wx0 * buf[q] + wx1 * buf[q + sy]) +
                       wy1 * (wx0 * buf[q + 1] + wx1 * buf[q + sy + 1]);
    }
    return;
  }
  b -= b2;
  const int i = b * 256 + tid;
  if (i >= t.nin) return;
  const double w3[3] = {-0.125, 0.75, 0.375};
  const int nxw = t.jkx[i] ? 3 : 1, nyw = t.jky[i] ? 3 : 1;
  double total = 0.0;
  for (int j = 0; j < nyw; ++j) {
    double row = 0.0;
    for (int k = 0; k < nxw; ++k) {
      const double wk = t.jkx[i] ? w3[k] : 1.0;
      row = row + wk * buf[t.jbase[i] + k * t.jsx[i] + j * t.jsy[i]];
    }
    const double wj = t.jky[i] ? w3[j] : 1.0;
    total = total + wj * row;
  }
  buf[t.jdst[i]] = total;
}

// A level pair's coarsening in one launch: a covered coarse cell becomes the mean of its
// four fine children, ((f00 + f10) + (f01 + f11)) / 4 (rows of 5 int64), and (the
// remaining workgroups) a coarse A_z node that coincides with a fine node takes the
// fine value (rows of 2 int64, mhd.py:981,991); cells and nodes are disjoint.
__global__ void __launch_bounds__(256) k_coarsen(double* __restrict__ buf,
                                                 const long long* __restrict__ arows, int na,
                                                 const long long* __restrict__ rrows, int nr) {
  const int b0 = nblocks(na);
  if ((int)blockIdx.x < b0) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= na) return;
    const long long* r = arows + 5 * (size_t)i;
    const long long coarse = r[0], pc = r[1], fine = r[2], fy = r[3], pf = r[4];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const long long q = fine + f * pf;
      buf[coarse + f * pc] = ((buf[q] + buf[q + fy]) + (buf[q + 1] + buf[q + fy + 1])) * 0.25;
    }
    return;
  }
  const int i = (blockIdx.x - b0) * 256 + threadIdx.x;
  if (i < nr) buf[rrows[2 * (size_t)i]] = buf[rrows[2 * (size_t)i + 1]];
}

// The outflow and wall ghosts of the patches on a non-periodic domain face
// (mhd.py:298-353,748-777), one workgroup per listed patch: rows {patch, faces,
// walls} with faces 1 x-low, 2 x-high, 4 y-low, 8 y-high and walls bit 0 (x walls),
// bit 1 (y walls).  Axis 0 first over the full haloed extent in y, then axis 1 over
// the full extent in x (so corner ghosts take the second rule on the first's
// values), each thread one line of a field: a cell field's ghosts copy the first
// interior cell (outflow) or mirror the interior with the sign -1 on the normal
// momentum (wall); A_z's ghost nodes extrapolate the edge node linearly (outflow) or
// are odd about the edge node (wall).
__global__ void k_domain_faces2d(double* __restrict__ buf, const PatchDesc* __restrict__ patches,
                                 const int* __restrict__ rows) {
  const int* r = rows + 3 * blockIdx.x;
  const PatchDesc pd = patches[r[0]];
  const int faces = r[1], walls = r[2];
  const int X = pd.nx + 4, Y = pd.ny + 4;
  const size_t PL = (size_t)X * Y;
  double* base = buf + pd.off;
  constexpr int h = HALO;
  for (int axis = 0; axis < 2; ++axis) {
    const bool lower = faces & (axis == 0 ? 1 : 4), upper = faces & (axis == 0 ? 2 : 8);
    if (lower || upper) {
      const bool wall = walls & (1 << axis);
      const int lc = axis == 0 ? X : Y, oc = axis == 0 ? Y : X;  // cell lengths
      const int nlines = 6 * oc + (oc + 1);                        // 6 fields, then A_z
      for (int t = threadIdx.x; t < nlines; t += blockDim.x) {
        const bool node = t >= 6 * oc;
        const int f = node ? 6 : t / oc, o = node ? t - 6 * oc : t % oc;
        const int len = node ? lc + 1 : lc;
        const size_t stride_other = axis == 0 ? 1 : (size_t)(node ? Y + 1 : Y);
        const size_t sa = axis == 0 ? (size_t)(node ? Y + 1 : Y) : 1;
        double* line = base + (size_t)f * PL + (size_t)o * stride_other;
        auto at = [&](int i) -> double& { return line[(size_t)i * sa]; };
        if (!node) {
          const double sign = (wall && f == 1 + axis) ? -1.0 : 1.0;
          if (wall) {
            if (lower)
              for (int k = 0; k < h; ++k) at(h - 1 - k) = sign * at(h + k);
            if (upper)
              for (int k = 0; k < h; ++k) at(len - h + k) = sign * at(len - h - 1 - k);
          } else {
            if (lower)
              for (int k = 0; k < h; ++k) at(k) = at(h);
            if (upper)
              for (int k = 0; k < h; ++k) at(len - h + k) = at(len - h - 1);
          }
        } else if (wall) {
          if (lower) {
            const double w = at(h);
            for (int k = 1; k <= h; ++k) at(h - k) = 2.0 * w - at(h + k);
          }
          if (upper) {
            const double w = at(len - h - 1);
            for (int k = 1; k <= h; ++k) at(len - h - 1 + k) = 2.0 * w - at(len - h - 1 - k);
          }
        } else {
          if (lower) {
            const double edge = at(h), inner = at(h + 1);
            for (int k = 1; k <= h; ++k) at(h - k) = edge + (double)k * (edge - inner);
          }
          if (upper) {
            const double edge = at(len - h - 1), inner = at(len - h - 2);
            for (int k = 1; k <= h; ++k) at(len - h - 1 + k) = edge + (double)k * (edge - inner);
          }
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace

extern "C" {

int wb_amr2d_domain_faces(double* buf, const void* patches, const int32_t* rows, int nrows,
                          void* stream) {
  if (nrows < 0 || (nrows > 0 && (!buf || !patches || !rows))) return WB_EINVAL;
  if (nrows == 0) return 0;
  hipLaunchKernelGGL(k_domain_faces2d, dim3((unsigned)nrows), dim3(256), 0, (hipStream_t)stream,
                     buf, (const PatchDesc*)patches, (const int*)rows);
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

int wb_amr2d_fill_level(double* buf, const long long* cdst, const long long* csrc,
                        const long long* cpd, const long long* cps, int ncells,
                        const long long* ndst, const long long* nsrc, int nnodes,
                        const long long* idst, const long long* ipd, const long long* ic00,
                        const long long* icy, const long long* ipc, const double* iwx,
                        const double* iwy, int nic, const long long* jdst,
                        const long long* jbase, const long long* jsx, const long long* jsy,
                        const long long* jkx, const long long* jky, int nin, void* stream) {
  if (ncells < 0 || nnodes < 0 || nic < 0 || nin < 0) return WB_EINVAL;
  const int n = nblocks(ncells) + nblocks(nnodes) + nblocks(nic) + nblocks(nin);
  if (n == 0) return 0;
  const FillLevel t{cdst, csrc, cpd, cps, ndst, nsrc, idst, ipd, ic00, icy, ipc, iwx, iwy,
                    jdst, jbase, jsx, jsy, jkx, jky, ncells, nnodes, nic, nin};
  hipLaunchKernelGGL(k_fill_level, dim3((unsigned)n), dim3(256), 0, (hipStream_t)stream, buf, t);
  return wb_last_error();
}

int wb_amr2d_coarsen(double* buf, const long long* arows, int na, const long long* rrows, int nr,
                     void* stream) {
  if (na < 0 || nr < 0) return WB_EINVAL;
  const int n = nblocks(na) + nblocks(nr);
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_coarsen, dim3((unsigned)n), dim3(256), 0, (hipStream_t)stream, buf, arows,
                     na, rrows, nr);
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
