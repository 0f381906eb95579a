// Real-space form of the default 4-point gradient.
//
// gradient_kernel(order=1) multiplies by i (8 sin w - sin 2w) / 6
// (kernels.py:63-66), which is exactly the Fourier multiplier of the central
// difference D f(c) = (8 (f(c+1) - f(c-1)) - (f(c+2) - f(c-2))) / 12 along the
// axis (with the forward transform's e^{-ikx} convention, f(c+1) <-> e^{ik}).
// So the three force meshes -IFFT[g_d P] of pm.py:54-56 equal -D_d phi with
// phi = IFFT[P]: one inverse FFT plus this stencil instead of three inverse
// FFTs.  D is antisymmetric (D^T = -D), which gives the adjoint below.
#include "common.hpp"

namespace jpm {

struct Mesh3 {
  uint32_t nx, ny, nz;
  FastDiv div_nz, div_ny;
};

// (i + d) mod n for |d| <= 2 and any n >= 1 (two corrections cover n = 1).
__device__ __forceinline__ uint32_t wrap_add(uint32_t i, int d, uint32_t n) {
  int r = int(i) + d;
  if (r < 0) r += int(n);
  if (r < 0) r += int(n);
  if (r >= int(n)) r -= int(n);
  if (r >= int(n)) r -= int(n);
  return uint32_t(r);
}

// out[d][c] = scale * -(D_d phi)(c), d = 0, 1, 2 (three meshes M apart).
__global__ void __launch_bounds__(256)
fd4_gradient_kernel(const float *__restrict__ phi, Mesh3 g, float scale, float *__restrict__ out) {
  const uint32_t M = g.nx * g.ny * g.nz;
  const float c8 = scale * (8.0f / 12.0f), c1 = scale * (1.0f / 12.0f);
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < M; e += gridDim.x * blockDim.x) {
    const uint32_t r = g.div_nz.div(e);
    const uint32_t k = e - r * g.nz;
    const uint32_t i = g.div_ny.div(r);
    const uint32_t j = r - i * g.ny;
    const uint32_t sx = g.ny * g.nz;
    const uint32_t row = e - k;            // (i, j, 0)
    const uint32_t plane = e - j * g.nz - k;  // (i, 0, k) minus k
    // x
    const float xp1 = phi[wrap_add(i, 1, g.nx) * sx + j * g.nz + k];
    const float xm1 = phi[wrap_add(i, -1, g.nx) * sx + j * g.nz + k];
    const float xp2 = phi[wrap_add(i, 2, g.nx) * sx + j * g.nz + k];
    const float xm2 = phi[wrap_add(i, -2, g.nx) * sx + j * g.nz + k];
    // y
    const float yp1 = phi[plane + wrap_add(j, 1, g.ny) * g.nz + k];
    const float ym1 = phi[plane + wrap_add(j, -1, g.ny) * g.nz + k];
    const float yp2 = phi[plane + wrap_add(j, 2, g.ny) * g.nz + k];
    const float ym2 = phi[plane + wrap_add(j, -2, g.ny) * g.nz + k];
    // z
    const float zp1 = phi[row + wrap_add(k, 1, g.nz)];
    const float zm1 = phi[row + wrap_add(k, -1, g.nz)];
    const float zp2 = phi[row + wrap_add(k, 2, g.nz)];
    const float zm2 = phi[row + wrap_add(k, -2, g.nz)];
    out[e] = -(c8 * (xp1 - xm1) - c1 * (xp2 - xm2));
    out[size_t(M) + e] = -(c8 * (yp1 - ym1) - c1 * (yp2 - ym2));
    out[2 * size_t(M) + e] = -(c8 * (zp1 - zm1) - c1 * (zp2 - zm2));
  }
}

// Adjoint of fd4_gradient_kernel: out = scale * sum_d (D_d g_d), since
// (-D_d)^T = D_d.
__global__ void __launch_bounds__(256)
fd4_divergence_kernel(const float *__restrict__ g3, size_t chan_stride, Mesh3 g, float scale, float *__restrict__ out,
                      uint32_t per) {
  // blockIdx.y: mesh of a batch of equal shapes (stride M); the three input
  // channels lie chan_stride floats apart
  const uint32_t M = g.nx * g.ny * g.nz;
  g3 += size_t(blockIdx.y) * M;
  out += size_t(blockIdx.y) * M;
  const float c8 = scale * (8.0f / 12.0f), c1 = scale * (1.0f / 12.0f);
  const float *gx = g3, *gy = g3 + chan_stride, *gz = g3 + 2 * chan_stride;
  // workgroup b takes the 256 cells of block (b mod 8) per + b / 8: the
  // dispatcher deals workgroups to the 8 XCDs in turn, so each XCD sweeps one
  // contiguous slab of x planes and the reads at x +- 1, 2 hit its L2
  const uint32_t blk = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  {
    const uint32_t e = blk * blockDim.x + threadIdx.x;
    if (e >= M) return;
    const uint32_t r = g.div_nz.div(e);
    const uint32_t k = e - r * g.nz;
    const uint32_t i = g.div_ny.div(r);
    const uint32_t j = r - i * g.ny;
    const uint32_t sx = g.ny * g.nz;
    const uint32_t row = e - k;
    const uint32_t plane = e - j * g.nz - k;
    const float dx = c8 * (gx[wrap_add(i, 1, g.nx) * sx + j * g.nz + k] - gx[wrap_add(i, -1, g.nx) * sx + j * g.nz + k]) -
                     c1 * (gx[wrap_add(i, 2, g.nx) * sx + j * g.nz + k] - gx[wrap_add(i, -2, g.nx) * sx + j * g.nz + k]);
    const float dy = c8 * (gy[plane + wrap_add(j, 1, g.ny) * g.nz + k] - gy[plane + wrap_add(j, -1, g.ny) * g.nz + k]) -
                     c1 * (gy[plane + wrap_add(j, 2, g.ny) * g.nz + k] - gy[plane + wrap_add(j, -2, g.ny) * g.nz + k]);
    const float dz = c8 * (gz[row + wrap_add(k, 1, g.nz)] - gz[row + wrap_add(k, -1, g.nz)]) -
                     c1 * (gz[row + wrap_add(k, 2, g.nz)] - gz[row + wrap_add(k, -2, g.nz)]);
    out[e] = (dx + dy) + dz;
  }
}

static Mesh3 make_mesh3(int nx, int ny, int nz) {
  Mesh3 g;
  g.nx = uint32_t(nx);
  g.ny = uint32_t(ny);
  g.nz = uint32_t(nz);
  g.div_nz = FastDiv(g.nz);
  g.div_ny = FastDiv(g.ny);
  return g;
}

int fd4_gradient(hipStream_t s, int nx, int ny, int nz, const float *phi, float scale, float *out3) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(phi && out3, "null pointer argument");
  const int64_t m = cells(nx, ny, nz);
  JPM_REQUIRE(3 * m <= (int64_t(1) << 32), "mesh too large");
  {
    JPM_TIMED(s, "fd4_gradient_kernel");
    hipLaunchKernelGGL(fd4_gradient_kernel, dim3(grid_for(m, 256, 256 * 128)), dim3(256), 0, s, phi,
                       make_mesh3(nx, ny, nz), scale, out3);
  }
  JPM_LAUNCHED("fd4_gradient_kernel");
  return JPM_OK;
}

int fd4_divergence_batch(hipStream_t s, int nx, int ny, int nz, int batch, const float *g3, size_t chan_stride,
                         float scale, float *out) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(g3 && out, "null pointer argument");
  JPM_REQUIRE(batch >= 1 && batch <= 65535, "batch of %d meshes", batch);
  const int64_t m = cells(nx, ny, nz);
  JPM_REQUIRE(3 * m <= (int64_t(1) << 32), "mesh too large");
  {
    JPM_TIMED(s, "fd4_divergence_kernel");
    const uint32_t blocks = uint32_t((m + 255) / 256), per = (blocks + 7) / 8;
    hipLaunchKernelGGL(fd4_divergence_kernel, dim3(8 * per, batch), dim3(256), 0, s, g3, chan_stride,
                       make_mesh3(nx, ny, nz), scale, out, per);
  }
  JPM_LAUNCHED("fd4_divergence_kernel");
  return JPM_OK;
}

int fd4_divergence(hipStream_t s, int nx, int ny, int nz, const float *g3, float scale, float *out) {
  return fd4_divergence_batch(s, nx, ny, nz, 1, g3, size_t(cells(nx, ny, nz)), scale, out);
}

}  // namespace jpm
