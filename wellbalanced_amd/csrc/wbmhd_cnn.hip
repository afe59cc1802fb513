// The learned refinement indicator: a 3-layer 3x3x3 CNN (8 -> 16 -> 16 -> 1 channels,
// ReLU) over a periodic 3D grid, BF16 operands on gfx950's matrix cores
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, and its training kernels.
//
// MFMA 16x16x32 operand layout (lane l, q = l / 16, r = l % 16): A[row r][k 8q..8q+7],
// B[k 8q..8q+7][col r], D[row 4q+i][col r], i = 0..3 (checked on the GPU).
//
// Activations are channel-last (voxel-major, z fastest, C bf16 per voxel).  A
// workgroup stages its input tile with a one-voxel halo (periodic wrap) in LDS once;
// its waves then read every tap's operands from LDS.
//
// * 16-channel outputs (layers 1 and 2, and layer 2's backward-data pass): implicit
//   GEMM with M = a run of 16 voxels along z, N = the 16 output channels, K = taps x
//   input channels, 32 per MFMA (two taps of 16 channels, or four of 8): 14 MFMAs per
//   run for 16 input channels, 7 for 8.  A tile is 4 x 8 runs (x, y), 8 per wave.
// * the 1-channel logit (layer 3): M = 16 voxels along z, N = 16 voxels along y, and
//   K = (dx, dz, y', ci) with the banded weights B[(y', ci)][j] = W[dx, y' - y_j, dz][ci]
//   (zero off the band): 81 MFMAs per 256 logits, no padded output channels.
// * the gradient through layer 3 (one channel in, 16 out): M = 16 voxels, N = the 16
//   channels, K = the 27 taps: one MFMA per run.
// * weight gradients: per tap M = input channels, N = output channels, K = 32 voxels
//   (two runs); layer 1 puts two taps of 8 channels in M; layer 3 has M = channels,
//   N = taps, K = voxels.
//
// Tap t = 9 (dx+1) + 3 (dy+1) + (dz+1) reads the input at v + (dx, dy, dz).
// Parameters: every layer is stored as FP32 masters W[t][ci][co] (27 x 16 x 16, the
// channels a layer lacks held at zero) and a bias of 16, so the weight-gradient
// kernels write a layer's gradient in the master layout.  The kernel layout is
// K[t][co][ci] (BF16, ci contiguous: one 16-byte load per lane); layer 2's
// backward-data pass is the forward kernel with K'[t][ci][co] = W[26 - t][ci][co];
// layer 3 reads row co = 0 of its forward layout.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wbmhd_common.hpp"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
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

// the periodic image of i on an axis of n voxels (any i: grids may be smaller than a tile)
__device__ __forceinline__ int wrap(int i, int n) {
  const int m = i % n;
  return m < 0 ? m + n : m;
}

__device__ __forceinline__ size_t vox(const Grid& g, int x, int y, int z) {
  return ((size_t)x * g.ny + y) * g.nz + z;
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 u) { return __builtin_bit_cast(bf16x8, u); }

__device__ __forceinline__ bf16x8 pack8(const unsigned short* h) {
  uint4 u;
  u.x = (unsigned)h[0] | ((unsigned)h[1] << 16);
  u.y = (unsigned)h[2] | ((unsigned)h[3] << 16);
  u.z = (unsigned)h[4] | ((unsigned)h[5] << 16);
  u.w = (unsigned)h[6] | ((unsigned)h[7] << 16);
  return as_bf16x8(u);
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

constexpr int TX = 4, TY = 8;  // a conv tile: 4 x 8 runs of 16 voxels along z
constexpr int HX = TX + 2, HY = TY + 2, HZ = 18;

// Stage the (TX+2) x (TY+2) x 18 input tile around runs (x0.., y0.., z0..) into LDS,
// C channels per voxel (C = 8 or 16; 16-byte chunks), periodic wrap.
template <int C, int SX, int SY>
__device__ __forceinline__ void stage_tile(unsigned short* tile, const unsigned short* in,
                                           const Grid& g, int x0, int y0, int z0) {
  constexpr int CH = C / 8, NV = SX * SY * HZ;
  for (int c = threadIdx.x; c < NV * CH; c += blockDim.x) {
    const int v = c / CH, part = c - v * CH;
    const int lz = v % HZ, t2 = v / HZ, ly = t2 % SY, lx = t2 / SY;
    const size_t src = vox(g, wrap(x0 + lx - 1, g.nx), wrap(y0 + ly - 1, g.ny), wrap(z0 + lz - 1, g.nz));
    *reinterpret_cast<uint4*>(tile + (size_t)v * C + part * 8) =
        *reinterpret_cast<const uint4*>(in + src * C + part * 8);
  }
}

// 16 output channels.  mode 0: out16 = bf16(relu(acc + bias)); mode 2: out16 =
// bf16(acc * (mask > 0)) (the ReLU derivative of the layer below).
template <int CIN>
__global__ void __launch_bounds__(256) k_conv16(const unsigned short* __restrict__ in,
                                                const unsigned short* __restrict__ wk,
                                                const float* __restrict__ bias, Grid g, int mode,
                                                const unsigned short* __restrict__ mask,
                                                unsigned short* __restrict__ out16) {
  __shared__ __attribute__((aligned(16))) unsigned short tile[HX * HY * HZ * CIN];
  constexpr int KB = CIN == 16 ? 14 : 7;
  const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY, z0 = blockIdx.z * 16;
  stage_tile<CIN, HX, HY>(tile, in, g, x0, y0, z0);
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15, wave = threadIdx.x >> 6;
  // B: K[t][co = r][ci0 .. ci0 + 7] of every K block
  bf16x8 b[KB];
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    const int t = CIN == 16 ? 2 * kb + (q >> 1) : 4 * kb + q;
    const int ci0 = CIN == 16 ? 8 * (q & 1) : 0;
    b[kb] = t < 27 ? as_bf16x8(*reinterpret_cast<const uint4*>(wk + ((size_t)t * 16 + r) * CIN + ci0))
                   : as_bf16x8(uint4{0u, 0u, 0u, 0u});
  }
  const float bco = bias ? bias[r] : 0.0f;
  __syncthreads();
  for (int run = wave; run < TX * TY; run += 4) {
    const int lx = run / TY, ly = run - lx * TY;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= g.nx || y >= g.ny) continue;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const int t = CIN == 16 ? 2 * kb + (q >> 1) : 4 * kb + q;
      const int ci0 = CIN == 16 ? 8 * (q & 1) : 0;
      bf16x8 a = as_bf16x8(uint4{0u, 0u, 0u, 0u});
      if (t < 27) {
        const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
        const int v = ((lx + 1 + dx) * HY + (ly + 1 + dy)) * HZ + (r + 1 + dz);
        a = as_bf16x8(*reinterpret_cast<const uint4*>(tile + v * CIN + ci0));
      }
      acc = mfma(a, b[kb], acc);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const size_t v = vox(g, x, y, z0 + 4 * q + i);
      if (mode == 0) {
        out16[v * 16 + r] = f2bf(fmaxf(acc[i] + bco, 0.0f));
      } else {
        const float m = bf2f(mask[v * 16 + r]) > 0.0f ? 1.0f : 0.0f;
        out16[v * 16 + r] = f2bf(acc[i] * m);
      }
    }
  }
}

// The logit (layer 3, 16 channels in, 1 out) of a tile of 4 x-columns x 16 y x 16 z,
// one wave per x-column: D[z][j] = sum over (dx, dz, y', ci) of in[x+dx, y', z+dz][ci]
// times the banded weight W3[dx, y' - (y0 + j), dz][ci] (zero unless |y' - y0 - j| <= 1).
constexpr int LX = 4, LY = 16;
__global__ void __launch_bounds__(64 * LX) k_conv_logit(const unsigned short* __restrict__ in,
                                                    const unsigned short* __restrict__ wk,
                                                    const float* __restrict__ bias, Grid g,
                                                    float* __restrict__ logit) {
  __shared__ __attribute__((aligned(16))) unsigned short tile[(LX + 2) * (LY + 2) * HZ * 16];
  __shared__ __attribute__((aligned(16))) unsigned short w3[27 * 16];
  const int x0 = blockIdx.x * LX, y0 = blockIdx.y * LY, z0 = blockIdx.z * 16;
  stage_tile<16, LX + 2, LY + 2>(tile, in, g, x0, y0, z0);
  for (int i = threadIdx.x; i < 27 * 16; i += blockDim.x) w3[i] = wk[(size_t)(i / 16) * 256 + (i % 16)];
  __syncthreads();
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15, wave = threadIdx.x >> 6;
  const int x = x0 + wave;
  const int ci0 = 8 * (q & 1);
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  const bf16x8 zero = as_bf16x8(uint4{0u, 0u, 0u, 0u});
  for (int dx = -1; dx <= 1; ++dx) {
    for (int dz = -1; dz <= 1; ++dz) {
#pragma unroll
      for (int bp = 0; bp < 9; ++bp) {
        const int yl = 2 * bp + (q >> 1);  // local y' of this lane's 8 k values, 0 .. 17
        const int v = ((wave + 1 + dx) * (LY + 2) + yl) * HZ + (r + 1 + dz);
        const bf16x8 a = as_bf16x8(*reinterpret_cast<const uint4*>(tile + v * 16 + ci0));
        const int dy = yl - 1 - r;  // y' - y for output column j = r
        const bf16x8 bw = (dy >= -1 && dy <= 1)
            ? as_bf16x8(*reinterpret_cast<const uint4*>(w3 + (9 * (dx + 1) + 3 * (dy + 1) + (dz + 1)) * 16 + ci0))
            : zero;
        acc = mfma(a, bw, acc);
      }
    }
  }
  const int y = y0 + r;
  if (x < g.nx && y < g.ny) {
    const float b3 = bias ? bias[0] : 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) logit[vox(g, x, y, z0 + 4 * q + i)] = acc[i] + b3;
  }
}

// The gradient through layer 3: g2[u][ci] = bf16((sum_t g3[u - s_t] W3[t][ci]) *
// (a2[u][ci] > 0)), g3 one bf16 per voxel; M = 16 voxels, N = 16 channels, K = 27 taps.
__global__ void __launch_bounds__(256) k_conv_back1(const unsigned short* __restrict__ g3,
                                                    const unsigned short* __restrict__ wk,
                                                    Grid g, const unsigned short* __restrict__ mask,
                                                    unsigned short* __restrict__ out16) {
  __shared__ unsigned short tile[HX * HY * HZ];
  const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY, z0 = blockIdx.z * 16;
  for (int v = threadIdx.x; v < HX * HY * HZ; v += blockDim.x) {
    const int lz = v % HZ, t2 = v / HZ, ly = t2 % HY, lx = t2 / HY;
    tile[v] = g3[vox(g, wrap(x0 + lx - 1, g.nx), wrap(y0 + ly - 1, g.ny), wrap(z0 + lz - 1, g.nz))];
  }
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15, wave = threadIdx.x >> 6;
  // B[t][ci = r] = W3[t][ci] for t = 8q .. 8q + 7 (zero beyond the 27 taps)
  unsigned short wb[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int t = 8 * q + j;
    wb[j] = t < 27 ? wk[(size_t)t * 256 + r] : (unsigned short)0;
  }
  const bf16x8 b = pack8(wb);
  __syncthreads();
  for (int run = wave; run < TX * TY; run += 4) {
    const int lx = run / TY, ly = run - lx * TY;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= g.nx || y >= g.ny) continue;
    unsigned short av[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {  // A[u = z0 + r][t] = g3[u - s_t]
      const int t = 8 * q + j;
      const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
      av[j] = t < 27 ? tile[((lx + 1 - dx) * HY + (ly + 1 - dy)) * HZ + (r + 1 - dz)]
                     : (unsigned short)0;
    }
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = mfma(pack8(av), b, acc);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const size_t v = vox(g, x, y, z0 + 4 * q + i);
      const float m = bf2f(mask[v * 16 + r]) > 0.0f ? 1.0f : 0.0f;
      out16[v * 16 + r] = f2bf(acc[i] * m);
    }
  }
}

// Weight gradient of a 16-output layer: dW[t][ci][co] = sum_v x[v + s_t][ci] g[v][co]
// (x has CIN channels), db[co] = sum_v g[v][co].  A persistent grid walks the conv
// tiles; per tile x (with halo) and g (16 channels) are staged in LDS, and each wave
// takes two runs at a time (K = 32 voxels).  Per workgroup the partial sums (27*256 +
// 16 floats, the master layout) go to `partial`, reduced in a fixed order by
// k_reduce_partials.  Taps 3 d, 3 d + 1, 3 d + 2 share (dx, dy) = d and their 8-voxel
// windows of a channel lie in 10 consecutive z values, read once from LDS: CIN = 16
// takes d = 0 .. 8 with M = the channels; CIN = 8 puts the groups d = 2p and 2p + 1 in
// rows 0-7 and 8-15 of M (p = 0 .. 4, d = 9 zero).
template <int CIN>
__global__ void __launch_bounds__(256) k_wgrad16(const unsigned short* __restrict__ x,
                                                 const unsigned short* __restrict__ gr, Grid g,
                                                 float* __restrict__ partial) {
  constexpr int NT = CIN == 16 ? 27 : 15;  // accumulators (M blocks of taps)
  constexpr int TXN = HX * HY * HZ * CIN, TGN = TX * TY * 16 * 16;
  constexpr int TILES = 2 * (TXN + TGN), RED = 4 * (27 * 256 + 16);
  // the tiles, and after the last tile the reduction of the waves' sums, in one buffer
  __shared__ __attribute__((aligned(16))) unsigned char smem[TILES > RED ? TILES : RED];
  unsigned short* tx = reinterpret_cast<unsigned short*>(smem);
  unsigned short* tg = tx + TXN;
  float* red = reinterpret_cast<float*>(smem);
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15, wave = threadIdx.x >> 6;
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float dbias = 0.0f;
  const int gx = (g.nx + TX - 1) / TX, gy = (g.ny + TY - 1) / TY, gz = g.nz / 16;
  const int ntiles = gx * gy * gz;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int bz = tile % gz, t2 = tile / gz, by = t2 % gy, bx = t2 / gy;
    const int x0 = bx * TX, y0 = by * TY, z0 = bz * 16;
    __syncthreads();  // the previous tile's reads are done
    stage_tile<CIN, HX, HY>(tx, x, g, x0, y0, z0);
    for (int c = threadIdx.x; c < TX * TY * 16 * 2; c += blockDim.x) {
      const int v = c >> 1, part = c & 1;
      const int lz = v % 16, t3 = v / 16, ly = t3 % TY, lx = t3 / TY;
      const int xx = wrap(x0 + lx, g.nx), yy = wrap(y0 + ly, g.ny);
      *reinterpret_cast<uint4*>(tg + (size_t)v * 16 + part * 8) =
          *reinterpret_cast<const uint4*>(gr + vox(g, xx, yy, z0 + lz) * 16 + part * 8);
    }
    __syncthreads();
    // run pairs (2p, 2p + 1) of the tile; lanes q = 0, 1 take run 2p, q = 2, 3 run 2p + 1
    for (int pair = wave; pair < TX * TY / 2; pair += 4) {
      const int run = 2 * pair + (q >> 1);
      const int lx = run / TY, ly = run - lx * TY;
      const bool valid = x0 + lx < g.nx && y0 + ly < g.ny;
      const int zb = 8 * (q & 1);  // this lane's 8 voxels: z0 + zb .. z0 + zb + 7
      unsigned short gv[8];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        gv[j] = valid ? tg[((lx * TY + ly) * 16 + zb + j) * 16 + r] : (unsigned short)0;
#pragma unroll
      for (int j = 0; j < 8; ++j) dbias += bf2f(gv[j]);
      const bf16x8 bfrag = pack8(gv);
      if constexpr (CIN == 16) {
        // taps 3 m', 3 m' + 1, 3 m' + 2 share (dx, dy): their 8-voxel windows of channel
        // ci = r lie in 10 consecutive z values, read once
#pragma unroll
        for (int dxy = 0; dxy < 9; ++dxy) {
          const int dx = dxy / 3 - 1, dy = dxy % 3 - 1;
          const int base = ((lx + 1 + dx) * HY + (ly + 1 + dy)) * HZ + zb;
          unsigned short xv[10];
#pragma unroll
          for (int j = 0; j < 10; ++j) xv[j] = tx[(base + j) * CIN + r];
#pragma unroll
          for (int dz = 0; dz < 3; ++dz) acc[3 * dxy + dz] = mfma(pack8(xv + dz), bfrag, acc[3 * dxy + dz]);
        }
      } else {
#pragma unroll
        for (int p = 0; p < 5; ++p) {
          const int dxy = 2 * p + (r >> 3), ci = r & 7;
          unsigned short xv[10];
          if (dxy < 9) {
            const int dx = dxy / 3 - 1, dy = dxy % 3 - 1;
            const int base = ((lx + 1 + dx) * HY + (ly + 1 + dy)) * HZ + zb;
#pragma unroll
            for (int j = 0; j < 10; ++j) xv[j] = tx[(base + j) * CIN + ci];
          } else {
#pragma unroll
            for (int j = 0; j < 10; ++j) xv[j] = 0;
          }
#pragma unroll
          for (int dz = 0; dz < 3; ++dz) acc[3 * p + dz] = mfma(pack8(xv + dz), bfrag, acc[3 * p + dz]);
        }
      }
    }
  }
  // lanes of one output channel co = r hold partial bias sums (q = 0..3)
  dbias += __shfl_xor(dbias, 16);
  dbias += __shfl_xor(dbias, 32);
  __syncthreads();
  for (int i = threadIdx.x; i < 27 * 256 + 16; i += blockDim.x) red[i] = 0.0f;
  __syncthreads();
  // D[row 4q + i][col co = r]: combine the waves in a fixed order
  for (int w = 0; w < 4; ++w) {
    if (wave == w) {
#pragma unroll
      for (int m = 0; m < NT; ++m)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = 4 * q + i;
          // CIN = 8: accumulator m = 3 p + dz holds group d = 2 p + (row >> 3)
          const int t = CIN == 16 ? m : 3 * (2 * (m / 3) + (row >> 3)) + m % 3;
          const int ci = CIN == 16 ? row : (row & 7);
          if (t < 27) red[t * 256 + ci * 16 + r] += acc[m][i];
        }
      if (q == 0) red[27 * 256 + r] += dbias;
    }
    __syncthreads();
  }
  float* out = partial + (size_t)blockIdx.x * (27 * 256 + 16);
  for (int i = threadIdx.x; i < 27 * 256 + 16; i += blockDim.x) out[i] = red[i];
}

// Weight gradient of layer 3: dW3[t][ci] = sum_u a2[u][ci] g3[u - s_t] (M = channels,
// N = taps in two blocks of 16, K = 32 voxels), db3 = sum g3; written in the master
// layout (co = 0; the other output channels zero).
__global__ void __launch_bounds__(256) k_wgrad1(const unsigned short* __restrict__ a2,
                                                const unsigned short* __restrict__ g3, Grid g,
                                                float* __restrict__ partial) {
  __shared__ __attribute__((aligned(16))) unsigned short ta[TX * TY * 16 * 16];
  __shared__ unsigned short tg[HX * HY * HZ];
  __shared__ float red[27 * 16 + 1];
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15, wave = threadIdx.x >> 6;
  f32x4 acc[2] = {f32x4{0.0f, 0.0f, 0.0f, 0.0f}, f32x4{0.0f, 0.0f, 0.0f, 0.0f}};
  float dbias = 0.0f;
  const int gx = (g.nx + TX - 1) / TX, gy = (g.ny + TY - 1) / TY, gz = g.nz / 16;
  const int ntiles = gx * gy * gz;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int bz = tile % gz, t2 = tile / gz, by = t2 % gy, bx = t2 / gy;
    const int x0 = bx * TX, y0 = by * TY, z0 = bz * 16;
    __syncthreads();
    for (int c = threadIdx.x; c < TX * TY * 16 * 2; c += blockDim.x) {
      const int v = c >> 1, part = c & 1;
      const int lz = v % 16, t3 = v / 16, ly = t3 % TY, lx = t3 / TY;
      const int xx = wrap(x0 + lx, g.nx), yy = wrap(y0 + ly, g.ny);
      *reinterpret_cast<uint4*>(ta + (size_t)v * 16 + part * 8) =
          *reinterpret_cast<const uint4*>(a2 + vox(g, xx, yy, z0 + lz) * 16 + part * 8);
    }
    for (int v = threadIdx.x; v < HX * HY * HZ; v += blockDim.x) {
      const int lz = v % HZ, t3 = v / HZ, ly = t3 % HY, lx = t3 / HY;
      tg[v] = g3[vox(g, wrap(x0 + lx - 1, g.nx), wrap(y0 + ly - 1, g.ny), wrap(z0 + lz - 1, g.nz))];
    }
    __syncthreads();
    for (int pair = wave; pair < TX * TY / 2; pair += 4) {
      const int run = 2 * pair + (q >> 1);
      const int lx = run / TY, ly = run - lx * TY;
      const bool valid = x0 + lx < g.nx && y0 + ly < g.ny;
      const int zb = 8 * (q & 1);
      unsigned short av[8];  // A[ci = r][k]: a2 of this lane's 8 voxels
#pragma unroll
      for (int j = 0; j < 8; ++j)
        av[j] = valid ? ta[((lx * TY + ly) * 16 + zb + j) * 16 + r] : (unsigned short)0;
      const bf16x8 afrag = pack8(av);
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) {  // B[k][t = 16 nb + r] = g3[u_k - s_t]
        const int t = 16 * nb + r;
        unsigned short gv[8];
        if (t < 27 && valid) {
          const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
          const int base = ((lx + 1 - dx) * HY + (ly + 1 - dy)) * HZ + (zb + 1 - dz);
#pragma unroll
          for (int j = 0; j < 8; ++j) gv[j] = tg[base + j];
        } else {
#pragma unroll
          for (int j = 0; j < 8; ++j) gv[j] = 0;
        }
        acc[nb] = mfma(afrag, pack8(gv), acc[nb]);
      }
      if (valid && r == 0) {
#pragma unroll
        for (int j = 0; j < 8; ++j) dbias += bf2f(tg[((lx + 1) * HY + (ly + 1)) * HZ + zb + 1 + j]);
      }
    }
  }
  dbias += __shfl_xor(dbias, 16);
  dbias += __shfl_xor(dbias, 32);
  __syncthreads();
  for (int i = threadIdx.x; i < 27 * 16 + 1; i += blockDim.x) red[i] = 0.0f;
  __syncthreads();
  for (int w = 0; w < 4; ++w) {
    if (wave == w) {
#pragma unroll
      for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int t = 16 * nb + r, ci = 4 * q + i;  // D[row ci][col t]
          if (t < 27) red[t * 16 + ci] += acc[nb][i];
        }
      if (lane == 0) red[27 * 16] += dbias;
    }
    __syncthreads();
  }
  float* out = partial + (size_t)blockIdx.x * (27 * 256 + 16);
  for (int i = threadIdx.x; i < 27 * 256 + 16; i += blockDim.x) {
    float v = 0.0f;
    if (i < 27 * 256 && (i & 15) == 0) v = red[(i >> 8) * 16 + ((i >> 4) & 15)];
    if (i == 27 * 256) v = red[27 * 16];
    out[i] = v;
  }
}

// Sum of the per-workgroup partials (width floats each) in a fixed order: thread (s,
// c) of a workgroup sums partials s, s + 16, ... of output c, then the 16 sums are
// added in order s = 0 .. 15 (deterministic).
__global__ void __launch_bounds__(256) k_reduce_partials(const float* __restrict__ partial, int nparts,
                                                         int width, float* __restrict__ out) {
  __shared__ float part[16][17];
  const int c = threadIdx.x & 15, s = threadIdx.x >> 4;
  const int i = blockIdx.x * 16 + c;
  float sum = 0.0f;
  if (i < width)
    for (int p = s; p < nparts; p += 16) sum += partial[(size_t)p * width + i];
  part[s][c] = sum;
  __syncthreads();
  if (s == 0 && i < width) {
    float total = 0.0f;
    for (int k = 0; k < 16; ++k) total += part[k][c];
    out[i] = total;
  }
}

// The BCE gradient of the logits (mean over n voxels) as one bf16 per voxel, and the
// loss: a fixed grid of BCE_GROUPS workgroups walks the voxels, each writes its sum to
// part[workgroup], and k_loss_sum adds those in a fixed tree (deterministic).
constexpr int BCE_GROUPS = 512;
__global__ void __launch_bounds__(256) k_bce(const float* __restrict__ logit,
                                             const float* __restrict__ label, size_t n,
                                             unsigned short* __restrict__ g3,
                                             float* __restrict__ part) {
  __shared__ float red[256];
  float l = 0.0f;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (size_t)gridDim.x * blockDim.x) {
    const float z = logit[v], y = label[v];
    const float s = 1.0f / (1.0f + __expf(-z));
    const float gv = (s - y) / (float)n;
    g3[v] = f2bf(gv);
    l += fmaxf(z, 0.0f) - z * y + log1pf(__expf(-fabsf(z)));
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(BCE_GROUPS) k_loss_sum(const float* __restrict__ part,
                                                         float* __restrict__ loss) {
  __shared__ float red[BCE_GROUPS];
  red[threadIdx.x] = part[threadIdx.x];
  __syncthreads();
  for (int s = BCE_GROUPS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0];
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
// BF16 kernel weights (uint16): forward layers 1..3, layer 2's backward-data pass, and a
// slot left unused (layer 3's backward pass reads its forward row co = 0).
constexpr int KW_F1 = 0, KW_F2 = KW_F1 + 27 * 16 * 8, KW_F3 = KW_F2 + WSIZE, KW_B2 = KW_F3 + WSIZE,
              KW_B3 = KW_B2 + WSIZE, KW_SIZE = KW_B3 + WSIZE;

inline dim3 conv_grid(const Grid& g) {
  return dim3((unsigned)((g.nx + TX - 1) / TX), (unsigned)((g.ny + TY - 1) / TY), (unsigned)(g.nz / 16));
}

template <int CIN>
void launch_conv16(hipStream_t s, const unsigned short* in, const unsigned short* wk,
                   const float* bias, Grid g, int mode, const unsigned short* mask,
                   unsigned short* out16) {
  hipLaunchKernelGGL(k_conv16<CIN>, conv_grid(g), dim3(256), 0, s, in, wk, bias, g, mode, mask,
                     out16);
}

void launch_logit(hipStream_t s, const unsigned short* in, const unsigned short* wk,
                  const float* bias, Grid g, float* logit) {
  const dim3 grid((unsigned)((g.nx + LX - 1) / LX), (unsigned)((g.ny + LY - 1) / LY),
                  (unsigned)(g.nz / 16));
  hipLaunchKernelGGL(k_conv_logit, grid, dim3(64 * LX), 0, s, in, wk, bias, g, logit);
}

void reduce(hipStream_t s, const float* partial, int workgroups, float* out) {
  hipLaunchKernelGGL(k_reduce_partials, dim3((LAYER + 15) / 16), dim3(256), 0, s, partial,
                     workgroups, LAYER, out);
}

template <int CIN>
void launch_wgrad16(hipStream_t s, int workgroups, const unsigned short* x,
                    const unsigned short* g16, Grid g, float* partial, float* out) {
  hipLaunchKernelGGL(k_wgrad16<CIN>, dim3((unsigned)workgroups), dim3(256), 0, s, x, g16, g,
                     partial);
  reduce(s, partial, workgroups, out);
}

}  // namespace

extern "C" {

int wb_cnn_conv(const uint16_t* in, int cin, const uint16_t* wk, const float* bias, int64_t nx,
                int64_t ny, int64_t nz, int mode, const uint16_t* mask, uint16_t* out16,
                float* logit, int workgroups, void* stream) {
  if (!grid_ok(nx, ny, nz) || (cin != 8 && cin != 16) || mode < 0 || mode > 2 || workgroups < 1)
    return WB_EINVAL;
  if ((mode == 1 && (cin != 16 || !logit)) || (mode != 1 && !out16) || (mode == 2 && !mask))
    return WB_EINVAL;
  Grid g{(int)nx, (int)ny, (int)nz};
  hipStream_t s = (hipStream_t)stream;
  if (mode == 1)
    launch_logit(s, (const unsigned short*)in, (const unsigned short*)wk, bias, g, logit);
  else if (cin == 16)
    launch_conv16<16>(s, in, wk, bias, g, mode, mask, out16);
  else
    launch_conv16<8>(s, in, wk, bias, g, mode, mask, out16);
  return wb_last_error();
}

int wb_cnn_wgrad(const uint16_t* x, int cin, const uint16_t* g16, int64_t nx, int64_t ny,
                 int64_t nz, float* partial, int workgroups, float* out, void* stream) {
  if (!grid_ok(nx, ny, nz) || (cin != 8 && cin != 16) || workgroups < 1) return WB_EINVAL;
  Grid g{(int)nx, (int)ny, (int)nz};
  hipStream_t s = (hipStream_t)stream;
  if (cin == 16)
    launch_wgrad16<16>(s, workgroups, x, g16, g, partial, out);
  else
    launch_wgrad16<8>(s, workgroups, x, g16, g, partial, out);
  return wb_last_error();
}

int wb_cnn_prep(const float* params, uint16_t* kw, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_prep_weights, dim3(27), dim3(256), 0, s, params, 8, kw + KW_F1,
                     (unsigned short*)nullptr);
  hipLaunchKernelGGL(k_prep_weights, dim3(27), dim3(256), 0, s, params + LAYER, 16, kw + KW_F2,
                     kw + KW_B2);
  hipLaunchKernelGGL(k_prep_weights, dim3(27), dim3(256), 0, s, params + 2 * LAYER, 16,
                     kw + KW_F3, (unsigned short*)nullptr);
  return wb_last_error();
}

int wb_cnn_forward(const uint16_t* x, const float* params, const uint16_t* kw, int64_t nx,
                   int64_t ny, int64_t nz, uint16_t* a1, uint16_t* a2, float* logit,
                   int workgroups, void* stream) {
  if (!grid_ok(nx, ny, nz) || workgroups < 1) return WB_EINVAL;
  Grid g{(int)nx, (int)ny, (int)nz};
  hipStream_t s = (hipStream_t)stream;
  launch_conv16<8>(s, x, kw + KW_F1, params + WSIZE, g, 0, nullptr, a1);
  launch_conv16<16>(s, a1, kw + KW_F2, params + LAYER + WSIZE, g, 0, nullptr, a2);
  launch_logit(s, a2, kw + KW_F3, params + 2 * LAYER + WSIZE, g, logit);
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
  // the per-workgroup losses go to `partial` (at least 6928 floats) before its use below
  hipLaunchKernelGGL(k_bce, dim3(BCE_GROUPS), dim3(256), 0, s, logit, label, n, g3, partial);
  hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(BCE_GROUPS), 0, s, partial, loss);
  hipLaunchKernelGGL(k_wgrad1, dim3((unsigned)workgroups), dim3(256), 0, s, a2, g3, g, partial);
  reduce(s, partial, workgroups, grad + 2 * LAYER);
  hipLaunchKernelGGL(k_conv_back1, conv_grid(g), dim3(256), 0, s, g3, kw + KW_F3, g, a2, g2);
  launch_wgrad16<16>(s, workgroups, a1, g2, g, partial, grad + LAYER);
  launch_conv16<16>(s, g2, kw + KW_B2, nullptr, g, 2, a1, g1);
  launch_wgrad16<8>(s, workgroups, x, g1, g, partial, grad);
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
