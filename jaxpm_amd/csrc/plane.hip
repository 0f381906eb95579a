// Two-dimensional painting for lens planes: the weighted 2-D CIC paint
// (painting.py:131-158), its adjoint, the density plane of a slab of the box
// (lensing.py:11-44) with its adjoint, and the Gaussian filter of
// utils.gaussian_smoothing (utils.py:208-222) applied between the library's
// 2-D transforms.
//
// A plane is small next to the particle arrays (1024^2 cells are 4 MB, 256^3
// particles 201 MB), so the paints stream the particles once and add their
// four deposits with float atomics, which L2 absorbs; the density plane
// reads only the particles and skips those outside the slab, whose deposits
// are zeros.
#include "tiles.hpp"

namespace jpm {

int fft_r2c(hipStream_t s, int nx, int ny, int nz, int batch, const float *in, void *out);
int fft_c2r(hipStream_t s, int nx, int ny, int nz, int batch, void *in, float *out);

// The four corners of a 2-D CIC deposit in the order of painting.py:141
// ((0,0), (1,0), (0,1), (1,1)): weights (1 - |x - cx|)(1 - |y - cy|)
// (painting.py:143-144), derivative factors -sign(x - cx) with sign(0) = +1
// (the derivative jax.grad takes of abs at 0), and cells int32(c) mod n
// (painting.py:148-150).
struct Cic2 {
  uint32_t idx[4];
  float w[4], dx[4], dy[4];
};

__device__ __forceinline__ void cic2(float x, float y, int nx, int ny, Cic2 &c) {
  const float fx = floorf(x), fy = floorf(y);
#pragma unroll
  for (int o = 0; o < 4; ++o) {
    const float cx = fx + float(o & 1), cy = fy + float(o >> 1);
    const float ex = x - cx, ey = y - cy;
    const float kx = 1.0f - fabsf(ex), ky = 1.0f - fabsf(ey);
    c.w[o] = kx * ky;
    c.dx[o] = (ex >= 0.0f ? -1.0f : 1.0f) * ky;
    c.dy[o] = kx * (ey >= 0.0f ? -1.0f : 1.0f);
    c.idx[o] = uint32_t(wrapi(cell_int(cx), nx)) * uint32_t(ny) + uint32_t(wrapi(cell_int(cy), ny));
  }
}

// mesh[nx][ny] += the deposits of particles (pos[p * stride], pos[p * stride
// + 1]), each times weight[p] when weight is given.
__global__ void __launch_bounds__(256)
paint2d_kernel(const float *__restrict__ pos, uint32_t n, uint32_t stride, const float *__restrict__ weight, int nx,
               int ny, float *__restrict__ mesh) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    Cic2 c;
    cic2(pos[size_t(p) * stride], pos[size_t(p) * stride + 1], nx, ny, c);
    const float wp = weight ? weight[p] : 1.0f;
#pragma unroll
    for (int o = 0; o < 4; ++o) atomicAdd(&mesh[c.idx[o]], weight ? c.w[o] * wp : c.w[o]);
  }
}

// The adjoint of paint2d for a mesh cotangent g: gpos[p] (2 floats at
// stride gstride) = w_p sum_o g[idx_o] dW_o/dxy, gweight[p] = sum_o g[idx_o]
// W_o.  Either output may be NULL.
__global__ void __launch_bounds__(256)
paint2d_vjp_kernel(const float *__restrict__ pos, uint32_t n, uint32_t stride, const float *__restrict__ weight, int nx,
                   int ny, const float *__restrict__ g, float *__restrict__ gpos, uint32_t gstride,
                   float *__restrict__ gweight) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    Cic2 c;
    cic2(pos[size_t(p) * stride], pos[size_t(p) * stride + 1], nx, ny, c);
    float sx = 0.0f, sy = 0.0f, sw = 0.0f;
#pragma unroll
    for (int o = 0; o < 4; ++o) {
      const float v = g[c.idx[o]];
      sx += v * c.dx[o];
      sy += v * c.dy[o];
      sw += v * c.w[o];
    }
    if (gpos) {
      const float wp = weight ? weight[p] : 1.0f;
      gpos[size_t(p) * gstride] = sx * wp;
      gpos[size_t(p) * gstride + 1] = sy * wp;
    }
    if (gweight) gweight[p] = sw;
  }
}

// The plane coordinates of lensing.py:26-29: jnp.mod(x, box_nx) /
// box_nx * res (both axes are wrapped and scaled by box_nx there).
__device__ __forceinline__ float plane_coord(float x, float bnx, float res) { return jmod(x, bnx) / bnx * res; }

// Density plane of the particles pos[n][3] with lo < z <= hi (the weight of
// lensing.py:32-33; the other particles' deposits are zeros and are
// skipped) into plane[res][res].
__global__ void __launch_bounds__(256)
density_plane_kernel(const float *__restrict__ pos, uint32_t n, float bnx, float lo, float hi, float res, int ires,
                     float *__restrict__ plane) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const float d = pos[3 * size_t(p) + 2];
    if (!(d > lo && d <= hi)) continue;
    Cic2 c;
    cic2(plane_coord(pos[3 * size_t(p)], bnx, res), plane_coord(pos[3 * size_t(p) + 1], bnx, res), ires, ires, c);
#pragma unroll
    for (int o = 0; o < 4; ++o) atomicAdd(&plane[c.idx[o]], c.w[o]);
  }
}

// Adjoint of the density plane for a plane cotangent g (already divided by
// the normalisation): gpos[p] = (res / box_nx) sum_o g[idx_o] dW_o/dxy for
// the particles in the slab (the mod has derivative 1, the selection 0),
// 0 for the others and for z.
__global__ void __launch_bounds__(256)
density_plane_vjp_kernel(const float *__restrict__ pos, uint32_t n, float bnx, float lo, float hi, float res, int ires,
                         const float *__restrict__ g, float *__restrict__ gpos) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    float *o3 = gpos + 3 * size_t(p);
    const float d = pos[3 * size_t(p) + 2];
    float sx = 0.0f, sy = 0.0f;
    if (d > lo && d <= hi) {
      Cic2 c;
      cic2(plane_coord(pos[3 * size_t(p)], bnx, res), plane_coord(pos[3 * size_t(p) + 1], bnx, res), ires, ires, c);
#pragma unroll
      for (int o = 0; o < 4; ++o) {
        const float v = g[c.idx[o]];
        sx += v * c.dx[o];
        sy += v * c.dy[o];
      }
      sx = sx * res / bnx;
      sy = sy * res / bnx;
    }
    o3[0] = sx;
    o3[1] = sy;
    o3[2] = 0.0f;
  }
}

__global__ void __launch_bounds__(256) plane_div_kernel(float *__restrict__ a, uint32_t m, float d) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) a[i] = a[i] / d;
}

__global__ void __launch_bounds__(256) plane_zero_kernel(float *__restrict__ a, uint32_t m) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) a[i] = 0.0f;
}

// fftfreq(n)[i] in float32: i / n for i < (n + 1) / 2, (i - n) / n above.
__device__ __forceinline__ float fftfreq(int i, int n) { return float(i < (n + 1) / 2 ? i : i - n) / float(n); }

// The Gaussian filter of utils.py:213-220 on the half spectrum of an
// [n0][n1] image: exp(-(|k| / s)^2 / 2) with |k| the norm of the two
// fftfreq wave numbers and s = 1 / (2 pi sigma) (norm.pdf(k, 0, s) /
// norm.pdf(0, 0, s)), times 1 / (n0 n1), the normalisation of ifft2.
__global__ void __launch_bounds__(256)
gaussian_filter_kernel(float2 *__restrict__ spec, int n0, int n1, float sd, float inv_m) {
  const int n1h = n1 / 2 + 1;
  const uint32_t H = uint32_t(n0) * uint32_t(n1h);
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < H; e += gridDim.x * blockDim.x) {
    const int a = int(e / uint32_t(n1h)), b = int(e % uint32_t(n1h));
    const float u = fftfreq(a, n0), v = fftfreq(b, n1);
    const float k = sqrtf(u * u + v * v);
    const float y = k / sd;
    const float f = expf(-(y * y) * 0.5f) * inv_m;
    const float2 z = spec[e];
    spec[e] = make_float2(z.x * f, z.y * f);
  }
}

int paint_2d(hipStream_t s, int nx, int ny, const float *pos, int64_t n, int stride, const float *weight,
             float *mesh) {
  JPM_REQUIRE(nx > 0 && ny > 0 && int64_t(nx) * ny <= kMaxCells, "plane shape (%d,%d) out of range", nx, ny);
  JPM_REQUIRE(n >= 0 && n <= int64_t(0xffffffffu), "particle count %lld out of range", (long long)n);
  JPM_REQUIRE(stride >= 2, "position stride %d must be at least 2", stride);
  if (n == 0) return JPM_OK;
  JPM_REQUIRE(pos && mesh, "null pointer argument");
  JPM_TIMED(s, "paint2d_kernel");
  hipLaunchKernelGGL(paint2d_kernel, dim3(grid_for(n, 256, 256 * 256)), dim3(256), 0, s, pos, uint32_t(n),
                     uint32_t(stride), weight, nx, ny, mesh);
  JPM_LAUNCHED("paint2d_kernel");
  return JPM_OK;
}

int paint_2d_vjp(hipStream_t s, int nx, int ny, const float *pos, int64_t n, int stride, const float *weight,
                 const float *g, float *gpos, int gstride, float *gweight) {
  JPM_REQUIRE(nx > 0 && ny > 0 && int64_t(nx) * ny <= kMaxCells, "plane shape (%d,%d) out of range", nx, ny);
  JPM_REQUIRE(n >= 0 && n <= int64_t(0xffffffffu), "particle count %lld out of range", (long long)n);
  JPM_REQUIRE(stride >= 2 && (!gpos || gstride >= 2), "strides must be at least 2");
  if (n == 0) return JPM_OK;
  JPM_REQUIRE(pos && g && (gpos || gweight), "null pointer argument");
  hipLaunchKernelGGL(paint2d_vjp_kernel, dim3(grid_for(n, 256, 256 * 256)), dim3(256), 0, s, pos, uint32_t(n),
                     uint32_t(stride), weight, nx, ny, g, gpos, uint32_t(gstride), gweight);
  JPM_LAUNCHED("paint2d_vjp_kernel");
  return JPM_OK;
}

// The slab bounds and the normalisation of lensing.py:32-40, evaluated as
// there: center -+ width / 2 and (box_nx / res)(box_ny / res) width in
// double precision, then rounded to float32 for the float32 comparisons and
// division.
struct PlaneArgs {
  float lo, hi, norm;
};

static PlaneArgs plane_args(int box_nx, int box_ny, double center, double width, int res) {
  PlaneArgs a;
  a.lo = float(center - width / 2.0);
  a.hi = float(center + width / 2.0);
  a.norm = float((double(box_nx) / double(res)) * (double(box_ny) / double(res)) * width);
  return a;
}

int density_plane(hipStream_t s, const float *pos, int64_t n, int box_nx, int box_ny, double center, double width,
                  int res, float *plane) {
  JPM_REQUIRE(box_nx > 0 && box_ny > 0, "box shape must be positive");
  JPM_REQUIRE(res > 0 && int64_t(res) * res <= kMaxCells, "plane resolution %d out of range", res);
  JPM_REQUIRE(n >= 0 && n <= int64_t(0xffffffffu), "particle count %lld out of range", (long long)n);
  JPM_REQUIRE(plane && (n == 0 || pos), "null pointer argument");
  const PlaneArgs a = plane_args(box_nx, box_ny, center, width, res);
  const uint32_t m = uint32_t(res) * uint32_t(res);
  JPM_TIMED(s, "density_plane");
  hipLaunchKernelGGL(plane_zero_kernel, dim3(grid_for(m, 256, 4096)), dim3(256), 0, s, plane, m);
  JPM_LAUNCHED("plane_zero_kernel");
  if (n > 0) {
    hipLaunchKernelGGL(density_plane_kernel, dim3(grid_for(n, 256, 256 * 256)), dim3(256), 0, s, pos, uint32_t(n),
                       float(box_nx), a.lo, a.hi, float(res), res, plane);
    JPM_LAUNCHED("density_plane_kernel");
  }
  hipLaunchKernelGGL(plane_div_kernel, dim3(grid_for(m, 256, 4096)), dim3(256), 0, s, plane, m, a.norm);
  JPM_LAUNCHED("plane_div_kernel");
  return JPM_OK;
}

int density_plane_vjp(hipStream_t s, const float *pos, int64_t n, int box_nx, int box_ny, double center,
                      double width, int res, float *g, float *gpos) {
  JPM_REQUIRE(box_nx > 0 && box_ny > 0, "box shape must be positive");
  JPM_REQUIRE(res > 0 && int64_t(res) * res <= kMaxCells, "plane resolution %d out of range", res);
  JPM_REQUIRE(n >= 0 && n <= int64_t(0xffffffffu), "particle count %lld out of range", (long long)n);
  if (n == 0) return JPM_OK;
  JPM_REQUIRE(pos && g && gpos, "null pointer argument");
  const PlaneArgs a = plane_args(box_nx, box_ny, center, width, res);
  const uint32_t m = uint32_t(res) * uint32_t(res);
  // the cotangent of the division comes first (g is overwritten with g / norm)
  hipLaunchKernelGGL(plane_div_kernel, dim3(grid_for(m, 256, 4096)), dim3(256), 0, s, g, m, a.norm);
  JPM_LAUNCHED("plane_div_kernel");
  hipLaunchKernelGGL(density_plane_vjp_kernel, dim3(grid_for(n, 256, 256 * 256)), dim3(256), 0, s, pos, uint32_t(n),
                     float(box_nx), a.lo, a.hi, float(res), res, g, gpos);
  JPM_LAUNCHED("density_plane_vjp_kernel");
  return JPM_OK;
}

size_t gaussian_smoothing_ws(int n0, int n1) {
  if (n0 <= 0 || n1 <= 0) return 0;
  return size_t(n0) * size_t(n1 / 2 + 1) * sizeof(float2);
}

// out = Re ifft2(fft2(im) filter) for a real [n0][n1] image: an R2C of the
// image as a 1 x n0 x n1 mesh, the filter, a C2R.  in and out may be the
// same array.
int gaussian_smoothing(hipStream_t s, int n0, int n1, const float *im, float *out, double sigma, void *ws,
                       size_t ws_bytes) {
  JPM_REQUIRE(n0 > 0 && n1 > 0 && int64_t(n0) * n1 <= kMaxCells, "image shape (%d,%d) out of range", n0, n1);
  JPM_REQUIRE(sigma > 0.0, "smoothing scale must be positive");
  JPM_REQUIRE(im && out, "null pointer argument");
  const size_t need = gaussian_smoothing_ws(n0, n1);
  JPM_REQUIRE(ws && ws_bytes >= need, "gaussian smoothing needs %zu workspace bytes, got %zu", need, ws_bytes);
  JPM_TIMED(s, "gaussian_smoothing");
  int rc = fft_r2c(s, 1, n0, n1, 1, im, ws);
  if (rc) return rc;
  const double scale = 1.0 / (2.0 * M_PI * sigma);
  const uint32_t H = uint32_t(n0) * uint32_t(n1 / 2 + 1);
  hipLaunchKernelGGL(gaussian_filter_kernel, dim3(grid_for(H, 256, 4096)), dim3(256), 0, s,
                     static_cast<float2 *>(ws), n0, n1, float(scale), float(1.0 / (double(n0) * double(n1))));
  JPM_LAUNCHED("gaussian_filter_kernel");
  return fft_c2r(s, 1, n0, n1, 1, ws, out);
}

}  // namespace jpm
