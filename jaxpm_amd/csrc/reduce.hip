// Deterministic float64-accumulated dot products of float32 vectors: the
// scalar cotangents of the leapfrog coefficients (<lambda_v, F>, <lambda_x,
// v>) and of the LPT growth factors.  A fixed grid of blocks writes partial
// sums; one block adds them in order, so the result repeats bitwise.
#include "common.hpp"

namespace jpm {

constexpr int kDotBlock = 256;
constexpr int kDotBlocks = 1024;

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ double block_sum_d(double v) {
  __shared__ double part[kDotBlock / 64];
  v = wave_sum_d(v);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < int(blockDim.x >> 6); ++w) s += part[w];
  return s;
}

__global__ void __launch_bounds__(kDotBlock)
dot_partial_kernel(const float *__restrict__ x, const float *__restrict__ y, int64_t n, double *__restrict__ partial) {
  double s = 0.0;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x)
    s += double(x[e]) * double(y[e]);
  s = block_sum_d(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kDotBlock)
dot_final_kernel(const double *__restrict__ partial, int blocks, double *__restrict__ out) {
  double s = 0.0;
  for (int b = threadIdx.x; b < blocks; b += blockDim.x) s += partial[b];
  s = block_sum_d(s);
  if (threadIdx.x == 0) *out = s;
}

size_t dot_ws() { return size_t(kDotBlocks) * sizeof(double); }

int dot(hipStream_t s, int64_t n, const float *x, const float *y, double *out, void *ws, size_t ws_bytes) {
  JPM_REQUIRE(n >= 0, "dot: negative length");
  JPM_REQUIRE(out && (n == 0 || (x && y)), "dot: null pointer argument");
  JPM_REQUIRE(ws && ws_bytes >= dot_ws(), "dot: workspace of %zu bytes, need %zu", ws_bytes, dot_ws());
  const int blocks = int(grid_for(n, kDotBlock, kDotBlocks));
  double *partial = static_cast<double *>(ws);
  {
    JPM_TIMED(s, "dot_partial_kernel");
    hipLaunchKernelGGL(dot_partial_kernel, dim3(blocks), dim3(kDotBlock), 0, s, x, y, n, partial);
    JPM_LAUNCHED("dot_partial_kernel");
  }
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kDotBlock), 0, s, partial, blocks, out);
  JPM_LAUNCHED("dot_final_kernel");
  return JPM_OK;
}

}  // namespace jpm
