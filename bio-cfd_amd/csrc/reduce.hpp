// Wavefront and workgroup reductions shared by the kernels (wave64:
// shuffles within the wavefront, one LDS slot per wavefront, one store per
// workgroup).
#pragma once

#include <hip/hip_runtime.h>

namespace bcfd {

__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_xor(v, o, 64);
    v = v > w ? v : w;
  }
  return v;
}

// workgroup max of a non-negative value -> partials[workgroup]
__device__ inline void wg_max_store(double v, double *partials) {
  __shared__ double s_wave[16];
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
  const int nw = (blockDim.x * blockDim.y + 63) / 64;
  v = wave_max(v);
  if ((t & 63) == 0) s_wave[t >> 6] = v;
  __syncthreads();
  if (t == 0) {
    double m = 0.0;
    for (int q = 0; q < nw; ++q) m = m > s_wave[q] ? m : s_wave[q];
    partials[blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z)] = m;
  }
}

// one workgroup: the max over n partials, to *out (thread 0)
__device__ inline void block_max_partials(const double *partials, int64_t n, double *out) {
  __shared__ double s_m[16];
  double m = 0.0;
  for (int64_t q = threadIdx.x; q < n; q += blockDim.x) m = m > partials[q] ? m : partials[q];
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int q = 0; q < (int)(blockDim.x + 63) / 64; ++q) r = r > s_m[q] ? r : s_m[q];
    *out = r;
  }
}

}  // namespace bcfd
