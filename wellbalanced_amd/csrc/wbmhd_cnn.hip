// The learned refinement indicator: a 3-layer 3x3x3 CNN (8 -> 16 -> 16 -> 1 channels,
// ReLU) over a periodic 3D grid, BF16 operands on the matrix cores (MFMA
// v_mfma_f32_16x16x16_bf16) with FP32 accumulation, and its training kernels.
//
// Implicit GEMM: activations are channel-last (voxel-major, C bf16 per voxel); a
// wavefront takes a run of 16 consecutive voxels along z (the M rows) and the 16
// output channels (N), and walks K = taps x input channels: 27 MFMAs per run for 16
// input channels, 14 for 8 (two taps per K block).  MFMA 16x16x16 operand layout
// (lane l, q = l / 16): A[row l%16][k 4q..4q+3], B[k 4q..4q+3][col l%16],
// D[row 4q+i][col l%16], i = 0..3.
//
// Tap t = 9 (dx+1) + 3 (dy+1) + (dz+1) reads the input at v + (dx, dy, dz), wrapped
// periodically.  Parameters: every layer is stored as FP32 masters W[t][ci][co]
// (27 x 16 x 16, the channels a layer lacks held at zero) and a bias of 16, so the
// weight-gradient kernel writes a layer's gradient in the master layout.  The
// kernel layout is K[t][co][ci] (BF16, ci contiguous: one 8-byte load per lane);
// the backward-data pass is the forward kernel with K'[t][ci][co] = W[26 - t][ci][co].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wbmhd_common.hpp"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (unsigned short)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

struct Grid {
  int nx, ny, nz;
};

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

__device__ __forceinline__ size_t vox(const Grid& g, int x, int y, int z) {
  return ((size_t)x * g.ny + y) * g.nz + z;
}

// Forward convolution (and, with mode 2, the backward-data pass).
// mode 0: out16 = bf16(relu(acc + bias)); mode 1: logit = acc + bias (channel 0);
// mode 2: out16 = bf16(acc * (mask > 0)) (the ReLU derivative of the layer below).
template <int CIN>
__global__ void __launch_bounds__(256) k_conv_fwd(const unsigned short* __restrict__ in,
                                                  const unsigned short* __restrict__ wk,
                                                  const float* __restrict__ bias, Grid g, int mode,
                                                  const unsigned short* __restrict__ mask,
                                                  unsigned short* __restrict__ out16,
                                                  float* __restrict__ logit) {
  constexpr int KB = CIN == 16 ? 27 : 14;  // K blocks of 16
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const int zruns = g.nz / 16;
  const int runs = g.nx * g.ny * zruns;
  // B fragments of every K block: 4 bf16 of K[t][co = r][ci]
  bf16x4 b[KB];
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    int t, c0;
    if (CIN == 16) {
      t = kb;
      c0 = 4 * q;
    } else {
      t = 2 * kb + (q >> 1);
      c0 = 4 * (q & 1);
    }
    if (t < 27) {
      const unsigned short* w = wk + ((size_t)t * 16 + r) * CIN + c0;
      b[kb] = bf16x4{(short)w[0], (short)w[1], (short)w[2], (short)w[3]};
    } else {
      b[kb] = bf16x4{0, 0, 0, 0};
    }
  }
  const float bco = bias ? bias[r] : 0.0f;
  for (int run = wave; run < runs; run += nwaves) {
    const int zr = run % zruns, xy = run / zruns, y = xy % g.ny, x = xy / g.ny;
    const int z0 = zr * 16;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      int t, c0;
      if (CIN == 16) {
        t = kb;
        c0 = 4 * q;
      } else {
        t = 2 * kb + (q >> 1);
        c0 = 4 * (q & 1);
      }
      bf16x4 a = {0, 0, 0, 0};
      if (t < 27) {
        const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
        const size_t v = vox(g, wrap(x + dx, g.nx), wrap(y + dy, g.ny), wrap(z0 + r + dz, g.nz));
        const uint2 raw = *reinterpret_cast<const uint2*>(in + v * CIN + c0);
        a = bf16x4{(short)(raw.x & 0xffff), (short)(raw.x >> 16), (short)(raw.y & 0xffff),
                   (short)(raw.y >> 16)};
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[kb], acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const size_t v = vox(g, x, y, z0 + 4 * q + i);
      if (mode == 0) {
        out16[v * 16 + r] = f2bf(fmaxf(acc[i] + bco, 0.0f));
      } else if (mode == 1) {
        if (r == 0) logit[v] = acc[i] + bco;
      } else {
        const float m = bf2f(mask[v * 16 + r]) > 0.0f ? 1.0f : 0.0f;
        out16[v * 16 + r] = f2bf(acc[i] * m);
      }
    }
  }
}

// Weight gradient dW[t][ci][co] = sum_v x[v + t][ci] g[v][co] (x has CIN channels,
// g 16), and db[co] = sum_v g[v][co]: per workgroup partial sums (27*256 + 16
// floats) into `partial`, reduced in a fixed order by k_reduce_partials.
template <int CIN>
__global__ void __launch_bounds__(256) k_conv_wgrad(const unsigned short* __restrict__ x,
                                                    const unsigned short* __restrict__ gr,
                                                    Grid g, float* __restrict__ partial) {
  __shared__ float red[27 * 256 + 16];
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  const int wib = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int i = threadIdx.x; i < 27 * 256 + 16; i += blockDim.x) red[i] = 0.0f;
  __syncthreads();
  const int zruns = g.nz / 16;
  const int runs = g.nx * g.ny * zruns;
  const int wave = blockIdx.x * nw + wib, nwaves = gridDim.x * nw;
  f32x4 acc[27];
#pragma unroll
  for (int t = 0; t < 27; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float dbias = 0.0f;
  for (int run = wave; run < runs; run += nwaves) {
    const int zr = run % zruns, xy = run / zruns, y = xy % g.ny, xx = xy / g.ny;
    const int z0 = zr * 16;
    // B: g[v_k][co = r] for the voxels k = 4q .. 4q+3 of the run
    unsigned short gb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      gb[i] = gr[vox(g, xx, y, z0 + 4 * q + i) * 16 + r];
      dbias += bf2f(gb[i]);
    }
    const bf16x4 bfrag = {(short)gb[0], (short)gb[1], (short)gb[2], (short)gb[3]};
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
      const int xs = wrap(xx + dx, g.nx), ys = wrap(y + dy, g.ny);
      short av[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int zs = wrap(z0 + 4 * q + i + dz, g.nz);
        av[i] = r < CIN ? (short)x[vox(g, xs, ys, zs) * CIN + r] : (short)0;
      }
      const bf16x4 afrag = {av[0], av[1], av[2], av[3]};
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(afrag, bfrag, acc[t], 0, 0, 0);
    }
  }
  // the four lanes of one output channel hold partial bias sums
  dbias += __shfl_xor(dbias, 16);
  dbias += __shfl_xor(dbias, 32);
  // D[row ci = 4q + i][col co = r]: combine the waves of the workgroup in LDS
  for (int w = 0; w < nw; ++w) {
    if (wib == w) {
#pragma unroll
      for (int t = 0; t < 27; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) red[t * 256 + (4 * q + i) * 16 + r] += acc[t][i];
      if (q == 0) red[27 * 256 + r] += dbias;
    }
    __syncthreads();
  }
  float* out = partial + (size_t)blockIdx.x * (27 * 256 + 16);
  for (int i = threadIdx.x; i < 27 * 256 + 16; i += blockDim.x) out[i] = red[i];
}

// Sum of the per-workgroup partials in workgroup order (deterministic).
__global__ void k_reduce_partials(const float* __restrict__ partial, int nparts, int width,
                                  float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= width) return;
  float s = 0.0f;
  for (int p = 0; p < nparts; ++p) s += partial[(size_t)p * width + i];
  out[i] = s;
}

// The BCE gradient of the logits (mean over n voxels), as a 16-channel bf16 field
// with the gradient in channel 0, and the loss summed per workgroup.
__global__ void k_bce(const float* __restrict__ logit, const float* __restrict__ label, size_t n,
                      unsigned short* __restrict__ g16, float* __restrict__ loss) {
  __shared__ float red[256];
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (v < n) {
    const float z = logit[v], y = label[v];
    const float s = 1.0f / (1.0f + __expf(-z));
    const float gv = (s - y) / (float)n;
    g16[v * 16] = f2bf(gv);
#pragma unroll
    for (int c = 1; c < 16; ++c) g16[v * 16 + c] = 0;
    l = fmaxf(z, 0.0f) - z * y + log1pf(__expf(-fabsf(z)));
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

__global__ void k_zero_f32(float* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

// Adam on the FP32 masters (Kingma and Ba 2015).
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = gr[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 kernel layouts from the FP32 masters W[t][ci][co] (27 x 16 x 16): forward
// K[t][co][ci] for ci < cin (8 or 16) and, if back is not null, backward-data
// K'[t][ci][co] = W[26 - t][ci][co].
__global__ void k_prep_weights(const float* __restrict__ w, int cin, unsigned short* __restrict__ fwd,
                               unsigned short* __restrict__ back) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 27 * 256) return;
  const int t = i / 256, a = (i / 16) % 16, b = i % 16;
  // forward: co = a, ci = b
  if (b < cin) fwd[((size_t)t * 16 + a) * cin + b] = f2bf(w[((size_t)t * 16 + b) * 16 + a]);
  // backward-data: output channel a (the layer's ci), input channel b (its co)
  if (back) back[((size_t)t * 16 + a) * 16 + b] = f2bf(w[((size_t)(26 - t) * 16 + a) * 16 + b]);
}

struct PrimArgs {
  const double* state;
  int nx, ny, nz;
  double dx, dy, dz, b0x, b0y, b0z, gamma;
};

// The 8 primitives (rho, u, v, w, p, Bx, By, Bz) of interior cell id (z fastest) of
// a haloed 3D state, as mhd3d.py:475-495: B the mean of the two face fields.
__device__ void cell_prims(const PrimArgs& a, size_t id, double prim[8]) {
  const int X = a.nx + 4, Y = a.ny + 4, Z = a.nz + 4;
  const int k = id % a.nz + 2, t = id / a.nz, j = t % a.ny + 2, i = t / a.ny + 2;
  const size_t P = (size_t)X * Y * Z;
  const double* ax = a.state + 5 * P;
  const double* ay = ax + (size_t)X * (Y + 1) * (Z + 1);
  const double* az = ay + (size_t)(X + 1) * Y * (Z + 1);
  auto AX = [&](int p, int q, int r) { return ax[((size_t)p * (Y + 1) + q) * (Z + 1) + r]; };
  auto AY = [&](int p, int q, int r) { return ay[((size_t)p * Y + q) * (Z + 1) + r]; };
  auto AZ = [&](int p, int q, int r) { return az[((size_t)p * (Y + 1) + q) * Z + r]; };
  auto bxf = [&](int p) {
    return (AZ(p, j + 1, k) - AZ(p, j, k)) / a.dy - (AY(p, j, k + 1) - AY(p, j, k)) / a.dz + a.b0x;
  };
  auto byf = [&](int q) {
    return (AX(i, q, k + 1) - AX(i, q, k)) / a.dz - (AZ(i + 1, q, k) - AZ(i, q, k)) / a.dx + a.b0y;
  };
  auto bzf = [&](int r) {
    return (AY(i + 1, j, r) - AY(i, j, r)) / a.dx - (AX(i, j + 1, r) - AX(i, j, r)) / a.dy + a.b0z;
  };
  const size_t c = ((size_t)i * Y + j) * Z + k;
  const double bx = 0.5 * (bxf(i) + bxf(i + 1)), by = 0.5 * (byf(j) + byf(j + 1)),
               bz = 0.5 * (bzf(k) + bzf(k + 1));
  const double r = a.state[c];
  const double u = a.state[P + c] / r, v = a.state[2 * P + c] / r, w = a.state[3 * P + c] / r;
  prim[0] = r;
  prim[1] = u;
  prim[2] = v;
  prim[3] = w;
  prim[4] = (a.gamma - 1.0) * (a.state[4 * P + c] - 0.5 * r * (u * u + v * v + w * w) -
                               0.5 * (bx * bx + by * by + bz * bz));
  prim[5] = bx;
  prim[6] = by;
  prim[7] = bz;
}

// Network input: the primitives normalised (x - mean) * inv_std, bf16 channel-last.
__global__ void k_cnn_input(PrimArgs a, const float* __restrict__ norm, unsigned short* __restrict__ out) {
  const size_t n = (size_t)a.nx * a.ny * a.nz;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  double prim[8];
  cell_prims(a, id, prim);
#pragma unroll
  for (int ch = 0; ch < 8; ++ch) out[id * 8 + ch] = f2bf(((float)prim[ch] - norm[ch]) * norm[8 + ch]);
}

// Per-workgroup sums of every primitive and of its square (FP64), 16 per workgroup.
__global__ void __launch_bounds__(256) k_cnn_moments(PrimArgs a, double* __restrict__ partial) {
  __shared__ double red[16][256];
  const size_t n = (size_t)a.nx * a.ny * a.nz;
  double s[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) s[c] = 0.0;
  for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n;
       id += (size_t)gridDim.x * blockDim.x) {
    double prim[8];
    cell_prims(a, id, prim);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      s[c] += prim[c];
      s[8 + c] += prim[c] * prim[c];
    }
  }
#pragma unroll
  for (int c = 0; c < 16; ++c) red[c][threadIdx.x] = s[c];
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int c = 0; c < 16; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 16) partial[(size_t)blockIdx.x * 16 + threadIdx.x] = red[threadIdx.x][0];
}

// mean and 1 / std of every channel from the workgroup sums, summed in order.
__global__ void k_cnn_norm(const double* __restrict__ partial, int nparts, double n,
                           float* __restrict__ norm) {
  const int c = threadIdx.x;
  if (c >= 8) return;
  double s = 0.0, q = 0.0;
  for (int p = 0; p < nparts; ++p) {
    s += partial[(size_t)p * 16 + c];
    q += partial[(size_t)p * 16 + 8 + c];
  }
  const double mean = s / n;
  const double var = q / n - mean * mean;
  norm[c] = (float)mean;
  norm[8 + c] = var > 1e-24 ? (float)(1.0 / sqrt(var)) : 1.0f;
}

inline bool grid_ok(int64_t nx, int64_t ny, int64_t nz) {
  return nx > 0 && ny > 0 && nz >= 16 && nz % 16 == 0 && nx * ny * nz < ((int64_t)1 << 31);
}

// Flat parameter layout (floats): per layer L = 0, 1, 2 the weights W[27][16][16] at
// L * LAYER and the bias[16] after them.
constexpr int WSIZE = 27 * 256, LAYER = WSIZE + 16, NPARAMS = 3 * LAYER;
// BF16 kernel weights (uint16): forward layers 1..3, then backward-data of layers 2, 3.
constexpr int KW_F1 = 0, KW_F2 = KW_F1 + 27 * 16 * 8, KW_F3 = KW_F2 + WSIZE, KW_B2 = KW_F3 + WSIZE,
              KW_B3 = KW_B2 + WSIZE, KW_SIZE = KW_B3 + WSIZE;

template <int CIN>
void launch_conv(hipStream_t s, int workgroups, const unsigned short* in, const unsigned short* wk,
                 const float* bias, Grid g, int mode, const unsigned short* mask,
                 unsigned short* out16, float* logit) {
  hipLaunchKernelGGL(k_conv_fwd<CIN>, dim3((unsigned)workgroups), dim3(256), 0, s, in, wk, bias, g,
                     mode, mask, out16, logit);
}

template <int CIN>
void launch_wgrad(hipStream_t s, int workgroups, const unsigned short* x, const unsigned short* g16,
                  Grid g, float* partial, float* out) {
  hipLaunchKernelGGL(k_conv_wgrad<CIN>, dim3((unsigned)workgroups), dim3(256), 0, s, x, g16, g,
                     partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3((LAYER + 255) / 256), dim3(256), 0, s, partial,
                     workgroups, LAYER, out);
}

}  // namespace

extern "C" {

int wb_cnn_conv(const uint16_t* in, int cin, const uint16_t* wk, const float* bias, int64_t nx,
                int64_t ny, int64_t nz, int mode, const uint16_t* mask, uint16_t* out16,
                float* logit, int workgroups, void* stream) {
  if (!grid_ok(nx, ny, nz) || (cin != 8 && cin != 16) || mode < 0 || mode > 2 || workgroups < 1)
    return WB_EINVAL;
  Grid g{(int)nx, (int)ny, (int)nz};
  hipStream_t s = (hipStream_t)stream;
  if (cin == 16)
    launch_conv<16>(s, workgroups, in, wk, bias, g, mode, mask, out16, logit);
  else
    launch_conv<8>(s, workgroups, in, wk, bias, g, mode, mask, out16, logit);
  return wb_last_error();
}

int wb_cnn_wgrad(const uint16_t* x, int cin, const uint16_t* g16, int64_t nx, int64_t ny,
                 int64_t nz, float* partial, int workgroups, float* out, void* stream) {
  if (!grid_ok(nx, ny, nz) || (cin != 8 && cin != 16) || workgroups < 1) return WB_EINVAL;
  Grid g{(int)nx, (int)ny, (int)nz};
  hipStream_t s = (hipStream_t)stream;
  if (cin == 16)
    launch_wgrad<16>(s, workgroups, x, g16, g, partial, out);
  else
    launch_wgrad<8>(s, workgroups, x, g16, g, partial, out);
  return wb_last_error();
}

int wb_cnn_prep(const float* params, uint16_t* kw, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_prep_weights, dim3(27), dim3(256), 0, s, params, 8, kw + KW_F1,
                     (unsigned short*)nullptr);
  hipLaunchKernelGGL(k_prep_weights, dim3(27), dim3(256), 0, s, params + LAYER, 16, kw + KW_F2,
                     kw + KW_B2);
  hipLaunchKernelGGL(k_prep_weights, dim3(27), dim3(256), 0, s, params + 2 * LAYER, 16,
                     kw + KW_F3, kw + KW_B3);
  return wb_last_error();
}

int wb_cnn_forward(const uint16_t* x, const float* params, const uint16_t* kw, int64_t nx,
                   int64_t ny, int64_t nz, uint16_t* a1, uint16_t* a2, float* logit,
                   int workgroups, void* stream) {
  if (!grid_ok(nx, ny, nz) || workgroups < 1) return WB_EINVAL;
  Grid g{(int)nx, (int)ny, (int)nz};
  hipStream_t s = (hipStream_t)stream;
  launch_conv<8>(s, workgroups, x, kw + KW_F1, params + WSIZE, g, 0, nullptr, a1, nullptr);
  launch_conv<16>(s, workgroups, a1, kw + KW_F2, params + LAYER + WSIZE, g, 0, nullptr, a2, nullptr);
  launch_conv<16>(s, workgroups, a2, kw + KW_F3, params + 2 * LAYER + WSIZE, g, 1, nullptr, nullptr,
                  logit);
  return wb_last_error();
}

int wb_cnn_backward(const uint16_t* x, const uint16_t* a1, const uint16_t* a2, const float* logit,
                    const float* label, const uint16_t* kw, int64_t nx, int64_t ny, int64_t nz,
                    uint16_t* g3, uint16_t* g2, uint16_t* g1, float* partial, int workgroups,
                    float* grad, float* loss, void* stream) {
  if (!grid_ok(nx, ny, nz) || workgroups < 1) return WB_EINVAL;
  Grid g{(int)nx, (int)ny, (int)nz};
  hipStream_t s = (hipStream_t)stream;
  const size_t n = (size_t)nx * ny * nz;
  hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, s, loss, 1);
  hipLaunchKernelGGL(k_bce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, logit, label, n,
                     g3, loss);
  launch_wgrad<16>(s, workgroups, a2, g3, g, partial, grad + 2 * LAYER);
  launch_conv<16>(s, workgroups, g3, kw + KW_B3, nullptr, g, 2, a2, g2, nullptr);
  launch_wgrad<16>(s, workgroups, a1, g2, g, partial, grad + LAYER);
  launch_conv<16>(s, workgroups, g2, kw + KW_B2, nullptr, g, 2, a1, g1, nullptr);
  launch_wgrad<8>(s, workgroups, x, g1, g, partial, grad);
  return wb_last_error();
}

int wb_cnn_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2,
                float eps, int step, void* stream) {
  if (n <= 0 || step < 1) return WB_EINVAL;
  const float c1 = 1.0f - __builtin_powf(b1, (float)step), c2 = 1.0f - __builtin_powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, p, g, m, v,
                     n, lr, b1, b2, eps, c1, c2);
  return wb_last_error();
}

int wb_cnn_moments(const double* state, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                   double dz, double b0x, double b0y, double b0z, double gamma, double* partial,
                   int workgroups, float* norm, void* stream) {
  if (nx < 1 || ny < 1 || nz < 1 || workgroups < 1) return WB_EINVAL;
  PrimArgs a{state, (int)nx, (int)ny, (int)nz, dx, dy, dz, b0x, b0y, b0z, gamma};
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_cnn_moments, dim3(workgroups), dim3(256), 0, s, a, partial);
  hipLaunchKernelGGL(k_cnn_norm, dim3(1), dim3(64), 0, s, partial, workgroups,
                     (double)(nx * ny * nz), norm);
  return wb_last_error();
}

int wb_cnn_input(const double* state, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                 double dz, double b0x, double b0y, double b0z, double gamma, const float* norm,
                 uint16_t* out, void* stream) {
  if (!grid_ok(nx, ny, nz)) return WB_EINVAL;
  PrimArgs a{state, (int)nx, (int)ny, (int)nz, dx, dy, dz, b0x, b0y, b0z, gamma};
  const size_t n = (size_t)nx * ny * nz;
  hipLaunchKernelGGL(k_cnn_input, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, a, norm, out);
  return wb_last_error();
}

}  // extern "C"
