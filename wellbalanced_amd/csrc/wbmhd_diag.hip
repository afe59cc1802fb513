// Diagnostics of a 2D hierarchy on the device (MHD2D.leaf_values, totals, divergence
// and composite_divergence, carafe mhd.py:1059-1170): the leaf cells' primitive
// variables and corner current, FP64 integrals of the conserved fields by a fixed
// two-pass tree (the same bits on every run), and the largest |div B| and coarse/fine
// node mismatch by ordered-key maxima.  FP64, -ffp-contract=off.
//
// A leaf table holds, per leaf cell, its patch and the haloed local cell (a, b) as
// a * Y + b; the patches are the hierarchy's PatchDesc table.
#include "nonideal.hpp"
#include "wbmhd_common.hpp"

using namespace wb;

namespace {

constexpr int TOTAL_GROUPS = 240;  // workgroups of the first totals pass

// mhd.py:1078-1098: prim of the cell (mhd_primitive) and the mean of J_z at its four
// corners (_current), summed in the reference's order.
__global__ void k_leaf_values(const double* __restrict__ buf, const PatchDesc* __restrict__ patches,
                              const int* __restrict__ leaf, int nleaf, double b0x, double b0y,
                              double gamma, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nleaf) return;
  const PatchDesc pd = patches[leaf[2 * t]];
  const int Y = pd.ny + 4, X = pd.nx + 4, YN = Y + 1;
  const int c = leaf[2 * t + 1], i = c / Y, j = c % Y;
  const double* in = buf + pd.off;
  const size_t PL = (size_t)X * Y;
  const double* az = in + 6 * PL;
  auto A = [&](int a, int b) { return az[(size_t)a * YN + b]; };
  const double bx0 = (A(i, j + 1) - A(i, j)) / pd.dy + b0x;
  const double bx1 = (A(i + 1, j + 1) - A(i + 1, j)) / pd.dy + b0x;
  const double by0 = -(A(i + 1, j) - A(i, j)) / pd.dx + b0y;
  const double by1 = -(A(i + 1, j + 1) - A(i, j + 1)) / pd.dx + b0y;
  const double r = in[c];
  const double u = in[PL + c] / r, v = in[2 * PL + c] / r, w = in[3 * PL + c] / r;
  const double bx = 0.5 * (bx0 + bx1);
  const double by = 0.5 * (by0 + by1);
  const double b = in[5 * PL + c];
  const double p = (gamma - 1.0) * (in[4 * PL + c] - 0.5 * r * (u * u + v * v + w * w) -
                                    0.5 * (bx * bx + by * by + b * b));
  double jz = node_current_az(az, YN, i, j, pd.dx, pd.dy, b0x, b0y);
  jz = jz + node_current_az(az, YN, i, j + 1, pd.dx, pd.dy, b0x, b0y);
  jz = jz + node_current_az(az, YN, i + 1, j, pd.dx, pd.dy, b0x, b0y);
  jz = jz + node_current_az(az, YN, i + 1, j + 1, pd.dx, pd.dy, b0x, b0y);
  const double vals[9] = {r, u, v, w, p, bx, by, b, 0.25 * jz};
#pragma unroll
  for (int k = 0; k < 9; ++k) out[(size_t)k * nleaf + t] = vals[k];
}

// Tree sum of the 256 values of a workgroup in LDS (fixed order).
__device__ __forceinline__ double tree_sum(double* s) {
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + h];
    __syncthreads();
  }
  return s[0];
}

// First pass: workgroup g sums leaf cells g*256 + k*(G*256) + tid of field f
// (blockIdx.y) times the cell area into partial[f * G + g].
__global__ void __launch_bounds__(256) k_totals_partial(const double* __restrict__ buf,
                                                        const PatchDesc* __restrict__ patches,
                                                        const int* __restrict__ leaf, int nleaf,
                                                        double* __restrict__ partial) {
  __shared__ double s[256];
  const int f = blockIdx.y, G = gridDim.x;
  double acc = 0.0;
  for (int t = blockIdx.x * 256 + threadIdx.x; t < nleaf; t += G * 256) {
    const PatchDesc pd = patches[leaf[2 * t]];
    const size_t PL = (size_t)(pd.nx + 4) * (pd.ny + 4);
    acc = acc + buf[pd.off + f * PL + leaf[2 * t + 1]] * (pd.dx * pd.dy);
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  const double total = tree_sum(s);
  if (threadIdx.x == 0) partial[f * G + blockIdx.x] = total;
}

__global__ void __launch_bounds__(256) k_totals_final(const double* __restrict__ partial, int G,
                                                      double* __restrict__ out) {
  __shared__ double s[256];
  const int f = blockIdx.x;
  s[threadIdx.x] = (int)threadIdx.x < G ? partial[f * G + threadIdx.x] : 0.0;
  __syncthreads();
  const double total = tree_sum(s);
  if (threadIdx.x == 0) out[f] = total;
}

// mhd.py:1114-1131 over every cell of every patch (chunks of 256 cells: patch, first
// cell): the largest |div B| times min(dx, dy), as an ordered key.
__global__ void __launch_bounds__(256) k_divergence(const double* __restrict__ buf,
                                                    const PatchDesc* __restrict__ patches,
                                                    const int* __restrict__ chunks, int nchunks,
                                                    double b0x, double b0y,
                                                    unsigned long long* __restrict__ key) {
  double worst = 0.0;
  for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const PatchDesc pd = patches[chunks[2 * ch]];
    const int t = chunks[2 * ch + 1] + (int)threadIdx.x;
    if (t >= pd.nx * pd.ny) continue;
    const int X = pd.nx + 4, Y = pd.ny + 4, YN = Y + 1;
    const int i = t / pd.ny + 2, j = t % pd.ny + 2;
    const double* az = buf + pd.off + 6 * (size_t)X * Y;
    auto A = [&](int a, int b) { return az[(size_t)a * YN + b]; };
    const double bx0 = (A(i, j + 1) - A(i, j)) / pd.dy + b0x;
    const double bx1 = (A(i + 1, j + 1) - A(i + 1, j)) / pd.dy + b0x;
    const double by0 = -(A(i + 1, j) - A(i, j)) / pd.dx + b0y;
    const double by1 = -(A(i + 1, j + 1) - A(i, j + 1)) / pd.dx + b0y;
    const double div = (bx1 - bx0) / pd.dx + (by1 - by0) / pd.dy;
    const double h = pd.dx < pd.dy ? pd.dx : pd.dy;
    const double d = __builtin_fabs(div) * h;
    worst = d > worst ? d : worst;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(worst, off, 64);
    worst = o > worst ? o : worst;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(key, order_key(worst));
}

// The largest |coarse node - fine node| over restriction rows (pairs of offsets).
__global__ void k_node_mismatch(const double* __restrict__ buf, const long long* __restrict__ rows,
                                int n, unsigned long long* __restrict__ key) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (i < n) d = __builtin_fabs(buf[rows[2 * (size_t)i]] - buf[rows[2 * (size_t)i + 1]]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(d, off, 64);
    d = o > d ? o : d;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(key, order_key(d));
}

// The largest difference of A_z between neighbouring nodes of the patches in `chunks`
// (the finer levels' patches): the scale of composite_divergence (mhd.py:1149-1152).
__global__ void __launch_bounds__(256) k_node_scale(const double* __restrict__ buf,
                                                    const PatchDesc* __restrict__ patches,
                                                    const int* __restrict__ chunks, int nchunks,
                                                    unsigned long long* __restrict__ key) {
  double big = 0.0;
  for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const PatchDesc pd = patches[chunks[2 * ch]];
    const int t = chunks[2 * ch + 1] + (int)threadIdx.x;
    if (t >= pd.nx * pd.ny) continue;
    const int X = pd.nx + 4, Y = pd.ny + 4, YN = Y + 1;
    const int i = t / pd.ny + 2, j = t % pd.ny + 2;
    const double* az = buf + pd.off + 6 * (size_t)X * Y;
    auto A = [&](int a, int b) { return az[(size_t)a * YN + b]; };
    // the x-edge and y-edge from node (i, j); the last row / column adds its far edge
    double d = __builtin_fabs(A(i + 1, j) - A(i, j));
    const double e = __builtin_fabs(A(i, j + 1) - A(i, j));
    d = e > d ? e : d;
    if (j == pd.ny + 1) {
      const double g = __builtin_fabs(A(i + 1, j + 1) - A(i, j + 1));
      d = g > d ? g : d;
    }
    if (i == pd.nx + 1) {
      const double g = __builtin_fabs(A(i + 1, j + 1) - A(i + 1, j));
      d = g > d ? g : d;
    }
    big = d > big ? d : big;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(big, off, 64);
    big = o > big ? o : big;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(key, order_key(big));
}

__global__ void k_key_init(unsigned long long* key, int n) {
  if ((int)threadIdx.x < n) key[threadIdx.x] = order_key(0.0);
}
__global__ void k_key_finish(double* out, int n) {
  if ((int)threadIdx.x < n) {
    const unsigned long long k = ((const unsigned long long*)out)[threadIdx.x];
    out[threadIdx.x] = key_to_double(k);
  }
}

}  // namespace

extern "C" {

int wb_amr2d_leaf_values(const double* buf, const void* patches, const int32_t* leaf, int nleaf,
                         double b0x, double b0y, double gamma, double* out, void* stream) {
  if (nleaf < 0) return WB_EINVAL;
  if (nleaf == 0) return 0;
  hipLaunchKernelGGL(k_leaf_values, dim3((unsigned)((nleaf + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, buf, (const PatchDesc*)patches, (const int*)leaf, nleaf,
                     b0x, b0y, gamma, out);
  return wb_last_error();
}

int wb_amr2d_totals_workspace(void) { return 6 * TOTAL_GROUPS; }

int wb_amr2d_totals(const double* buf, const void* patches, const int32_t* leaf, int nleaf,
                    double* workspace, double* out, void* stream) {
  if (nleaf < 0) return WB_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_totals_partial, dim3(TOTAL_GROUPS, 6), dim3(256), 0, s, buf,
                     (const PatchDesc*)patches, (const int*)leaf, nleaf, workspace);
  hipLaunchKernelGGL(k_totals_final, dim3(6), dim3(256), 0, s, workspace, TOTAL_GROUPS, out);
  return wb_last_error();
}

int wb_amr2d_divergence(const double* buf, const void* patches, const int32_t* chunks,
                        int nchunks, double b0x, double b0y, const long long* rows, int nrows,
                        const int32_t* fine_chunks, int nfine, double* out, void* stream) {
  if (nchunks < 1 || nrows < 0 || nfine < 0) return WB_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  unsigned long long* keys = (unsigned long long*)out;
  hipLaunchKernelGGL(k_key_init, dim3(1), dim3(64), 0, s, keys, 3);
  const unsigned grid = (unsigned)(nchunks < 1024 ? nchunks : 1024);
  hipLaunchKernelGGL(k_divergence, dim3(grid), dim3(256), 0, s, buf, (const PatchDesc*)patches,
                     (const int*)chunks, nchunks, b0x, b0y, keys);
  if (nrows > 0)
    hipLaunchKernelGGL(k_node_mismatch, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, s,
                       buf, rows, nrows, keys + 1);
  if (nfine > 0)
    hipLaunchKernelGGL(k_node_scale, dim3((unsigned)(nfine < 1024 ? nfine : 1024)), dim3(256), 0,
                       s, buf, (const PatchDesc*)patches, (const int*)fine_chunks, nfine,
                       keys + 2);
  hipLaunchKernelGGL(k_key_finish, dim3(1), dim3(64), 0, s, out, 3);
  return wb_last_error();
}

}  // extern "C"
