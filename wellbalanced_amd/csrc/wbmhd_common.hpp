// Shared host/device helpers of the wbmhd library: error codes, the ordered-key
// reduction of the CFL rate and the smallest density, and small utility kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wbmhd.h"

static inline int wb_last_error() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : -(int)e;
}

namespace wb {

// A patch of a hierarchy buffer and a unit of work (one tile x segment of a patch).
struct PatchDesc {
  long long off;  // first double of the patch's state
  int nx, ny;
  double dx, dy;
  int walls;  // the patch's faces on domain walls: 1 x-low, 2 x-high, 4 y-low, 8 y-high
  int pad;
};
struct WorkItem {
  int patch, tile, seg;
};

// A double as an unsigned key whose integer order is the order of the doubles
// (sign-magnitude to two's-complement style flip), so atomicMin/atomicMax on the
// keys reduce doubles exactly and in any order.
__device__ __forceinline__ unsigned long long order_key(double d) {
  unsigned long long u = (unsigned long long)__double_as_longlong(d);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ inline double key_to_double(unsigned long long k) {
  unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  double d;
  __builtin_memcpy(&d, &u, sizeof d);
  return d;
}

// keys[0]: max rate (starts at key(0)); keys[1]: min density (starts at key(+inf)).
static __global__ void k_rate_init(unsigned long long* keys) {
  if (threadIdx.x == 0) {
    keys[0] = order_key(0.0);
    keys[1] = order_key(__builtin_inf());
  }
}

static __global__ void k_rate_finish(const unsigned long long* keys, double* out) {
  if (threadIdx.x == 0) {
    const unsigned long long k0 = keys[0], k1 = keys[1];
    out[0] = key_to_double(k0);
    out[1] = key_to_double(k1);
  }
}

static __global__ void k_zero_i32(int32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Wavefront max/min by lane shuffles, then one atomic per wavefront.  A rate that
// is not > 0 (including NaN) does not raise the maximum, as in
// mhd_rate_pyccel.py:78-81; a NaN density does not lower the minimum.
__device__ __forceinline__ void block_reduce_rate_rho(double rate, double r,
                                                      unsigned long long* keys) {
  if (!(rate > 0.0)) rate = 0.0;
  if (!(r == r)) r = __builtin_inf();
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(rate, off, 64);
    const double p = __shfl_xor(r, off, 64);
    rate = o > rate ? o : rate;
    r = p < r ? p : r;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&keys[0], order_key(rate));
    atomicMin(&keys[1], order_key(r));
  }
}

// Block-wide max rate / min density: wavefront shuffles, the waves combined in
// LDS, then ONE atomic pair per workgroup (blockDim a multiple of 64, <= 1024; every
// thread of the block must call it).
__device__ __forceinline__ void block_reduce_rate_rho_once(double rate, double r,
                                                           unsigned long long* keys) {
  __shared__ double s_rate[16], s_rho[16];
  if (!(rate > 0.0)) rate = 0.0;
  if (!(r == r)) r = __builtin_inf();
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(rate, off, 64);
    const double p = __shfl_xor(r, off, 64);
    rate = o > rate ? o : rate;
    r = p < r ? p : r;
  }
  const int wave = threadIdx.x >> 6, waves = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_rate[wave] = rate;
    s_rho[wave] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = s_rate[0], b = s_rho[0];
    for (int w = 1; w < waves; ++w) {
      a = s_rate[w] > a ? s_rate[w] : a;
      b = s_rho[w] < b ? s_rho[w] : b;
    }
    atomicMax(&keys[0], order_key(a));
    atomicMin(&keys[1], order_key(b));
  }
}

}  // namespace wb
