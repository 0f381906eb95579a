// Float64 CIC paint, CIC read and pm_forces: the reference's computations
// under jax_enable_x64 (its distributed-equivalence tests run in float64,
// tests/test_distributed_pm.py:30).  The corner maths is painting.py's and
// painting_utils.py's in double precision; deposits are summed with float64
// atomics (no fixed-point mode: a 64-bit fixed point would hold fewer
// significant bits than the double sums); the transforms are hipFFT D2Z /
// Z2D with the k-space multipliers of kernels.py evaluated in double.
#include "common.hpp"

namespace jpm {

namespace {

struct AxisD {
  int i0, i1;       // wrapped indices of the low / high corner
  double w0, w1;    // weights 1 - |x - c| (0 for a dropped relative corner)
};

__device__ __forceinline__ int cell_int_d(double f) {
  return (f >= -2147483648.0 && f < 2147483648.0) ? int(f) : 0;
}

__device__ __forceinline__ int wrap_d(int c, int n) {
  if (uint32_t(c) < uint32_t(n)) return c;
  int r = c % n;
  return r < 0 ? r + n : r;
}

// painting.py:21-37 in double: floor, corner f + o, weight 1 - |x - c|,
// int32 cast then periodic mod.
__device__ __forceinline__ AxisD axis_abs_d(double x, int n) {
  AxisD a;
  const double f = floor(x), f1 = f + 1.0;
  a.w0 = 1.0 - fabs(x - f);
  a.w1 = 1.0 - fabs(x - f1);
  a.i0 = wrap_d(cell_int_d(f), n);
  a.i1 = wrap_d(cell_int_d(f1), n);
  return a;
}

// painting_utils.py:48-69 in double: q = (p + o) mod L (jnp.remainder),
// floor, nd = p - q, nd -= rint(nd / L) L, weight 1 - |nd|; the corner is
// dropped (zero weight at a valid index) unless 0 <= q < L.
__device__ __forceinline__ void corner_rel_d(double p, double o, double L, int &idx, double &w) {
  double q = fmod(p + o, L);
  if (q != 0.0 && q < 0.0) q = q + L;
  q = floor(q);
  double nd = p - q;
  nd = nd - rint(nd / L) * L;
  const bool ok = (q >= 0.0) && (q < L);
  idx = ok ? int(q) : 0;
  w = ok ? 1.0 - fabs(nd) : 0.0;
}

struct GridD {
  int nx, ny, nz;
  FastDiv div_nz, div_ny;
};

// The 8 corners (flat index, weight) of particle p: absolute corner order
// 000,100,010,001,110,101,011,111 (painting.py:24-25) or, relative, bit d of
// the corner number along axis d (painting_utils.py:43-45).  A dropped
// relative corner has weight 0 at index 0.
template <bool REL>
__device__ __forceinline__ void corners_d(const double *__restrict__ pos, uint32_t p, const GridD &g, int64_t idx[8],
                                          double w[8]) {
  const double x = pos[3 * size_t(p)], y = pos[3 * size_t(p) + 1], z = pos[3 * size_t(p) + 2];
  AxisD ax, ay, az;
  if (REL) {
    const uint32_t t = g.div_nz.div(p);
    const uint32_t k = p - t * uint32_t(g.nz);
    const uint32_t i = g.div_ny.div(t);
    const uint32_t j = t - i * uint32_t(g.ny);
    corner_rel_d(double(i) + x, 0.0, double(g.nx), ax.i0, ax.w0);
    corner_rel_d(double(i) + x, 1.0, double(g.nx), ax.i1, ax.w1);
    corner_rel_d(double(j) + y, 0.0, double(g.ny), ay.i0, ay.w0);
    corner_rel_d(double(j) + y, 1.0, double(g.ny), ay.i1, ay.w1);
    corner_rel_d(double(k) + z, 0.0, double(g.nz), az.i0, az.w0);
    corner_rel_d(double(k) + z, 1.0, double(g.nz), az.i1, az.w1);
  } else {
    ax = axis_abs_d(x, g.nx);
    ay = axis_abs_d(y, g.ny);
    az = axis_abs_d(z, g.nz);
  }
  const int xs[8] = {0, 1, 0, 0, 1, 1, 0, 1};
  const int ys[8] = {0, 0, 1, 0, 1, 0, 1, 1};
  const int zs[8] = {0, 0, 0, 1, 0, 1, 1, 1};
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const int bx = REL ? (o & 1) : xs[o], by = REL ? ((o >> 1) & 1) : ys[o], bz = REL ? ((o >> 2) & 1) : zs[o];
    const int ix = bx ? ax.i1 : ax.i0, iy = by ? ay.i1 : ay.i0, iz = bz ? az.i1 : az.i0;
    w[o] = (bx ? ax.w1 : ax.w0) * (by ? ay.w1 : ay.w0) * (bz ? az.w1 : az.w0);
    idx[o] = (int64_t(ix) * g.ny + iy) * g.nz + iz;
  }
}

template <bool REL>
__global__ void __launch_bounds__(256)
paint_f64_kernel(const double *__restrict__ pos, const double *__restrict__ weight, double wscalar, uint32_t n,
                 GridD g, double *__restrict__ mesh) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    int64_t idx[8];
    double w[8];
    corners_d<REL>(pos, p, g, idx, w);
    const double wp = weight ? weight[p] : wscalar;
#pragma unroll
    for (int o = 0; o < 8; ++o) unsafeAtomicAdd(&mesh[idx[o]], wp * w[o]);
  }
}

template <bool REL>
__global__ void __launch_bounds__(256)
read_f64_kernel(const double *__restrict__ mesh, int nc, int64_t cstride, const double *__restrict__ pos, uint32_t n,
                GridD g, double *__restrict__ out) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    int64_t idx[8];
    double w[8];
    corners_d<REL>(pos, p, g, idx, w);
    for (int c = 0; c < nc; ++c) {
      const double *m = mesh + c * cstride;
      double s = 0.0;
#pragma unroll
      for (int o = 0; o < 8; ++o) s += m[idx[o]] * w[o];
      out[size_t(p) * nc + c] = s;
    }
  }
}

__global__ void __launch_bounds__(256) zero_f64_kernel(double *__restrict__ p, int64_t m) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < m; i += int64_t(gridDim.x) * blockDim.x)
    p[i] = 0.0;
}

// k = 2 pi fftfreq(n)[i] (kernels.py:10-23) in double.
__device__ __forceinline__ double wavenumber(int i, int n) {
  const int f = i <= (n - 1) / 2 ? i : i - n;
  return 2.0 * M_PI * double(f) / double(n);
}

// sinc(t) = sin(pi t) / (pi t), 1 at 0 (jnp.sinc).
__device__ __forceinline__ double sinc_d(double t) {
  if (t == 0.0) return 1.0;
  const double a = M_PI * t;
  return sin(a) / a;
}

// The three force spectra of pm.py:49-56 from the half spectrum of the
// density: F_d = -g_d(k) * invlaplace(k) * longrange(k) * delta(k), with
// g_d the 4-point gradient i (8 sin w - sin 2w) / 6 (order 1) or i k with
// the len//2 entry zeroed (order 0, JPM_K_GRAD_EXACT), invlaplace -1/kk (0 at
// k = 0; with JPM_K_LAPL_FD, kk from k sinc(k / 2 pi)), longrange
// exp(-kk r^2) (kernels.py:41-115), and with JPM_K_DECONV the CIC
// compensation prod sinc(k / 2 pi)^-2 (kernels.py:118-136).
__global__ void __launch_bounds__(256)
force_f64_kernel(const double2 *__restrict__ dk, int nx, int ny, int nz, double r2, uint32_t kflags,
                 double2 *__restrict__ out3) {
  const int nzh = nz / 2 + 1;
  const int64_t H = int64_t(nx) * ny * nzh;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < H; e += int64_t(gridDim.x) * blockDim.x) {
    const int kz = int(e % nzh);
    const int64_t r = e / nzh;
    const int ky = int(r % ny), kx = int(r / ny);
    const double k[3] = {wavenumber(kx, nx), wavenumber(ky, ny), wavenumber(kz, nz)};
    const int idx[3] = {kx, ky, kz}, len[3] = {nx, ny, nz};
    double kk = 0.0, kk_plain = 0.0;
    for (int d = 0; d < 3; ++d) {
      kk_plain += k[d] * k[d];
      const double kd = (kflags & JPM_K_LAPL_FD) ? k[d] * sinc_d(k[d] / (2.0 * M_PI)) : k[d];
      kk += kd * kd;
    }
    double pot = kk == 0.0 ? 0.0 : -1.0 / kk;
    if (r2 != 0.0) pot *= exp(-kk_plain * r2);
    if (kflags & JPM_K_DECONV) {
      const double s = sinc_d(k[0] / (2.0 * M_PI)) * sinc_d(k[1] / (2.0 * M_PI)) * sinc_d(k[2] / (2.0 * M_PI));
      pot /= s * s;
    }
    const double2 v = dk[e];
    for (int d = 0; d < 3; ++d) {
      double g;   // gradient_kernel is i * g
      if (kflags & JPM_K_GRAD_EXACT)
        g = (idx[d] == len[d] / 2) ? 0.0 : k[d];
      else
        g = (8.0 * sin(k[d]) - sin(2.0 * k[d])) / 6.0;
      // -(i g) * pot * v
      const double a = -g * pot;
      out3[d * H + e] = make_double2(-a * v.y, a * v.x);
    }
  }
}

__global__ void __launch_bounds__(256) scale_f64_kernel(double *__restrict__ p, int64_t m, double s) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < m; i += int64_t(gridDim.x) * blockDim.x)
    p[i] *= s;
}

GridD make_grid_d(int nx, int ny, int nz) {
  GridD g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.div_nz = FastDiv(uint32_t(nz));
  g.div_ny = FastDiv(uint32_t(ny));
  return g;
}

}  // namespace

int fft_d2z(hipStream_t s, int nx, int ny, int nz, int batch, const double *in, void *out);
int fft_z2d(hipStream_t s, int nx, int ny, int nz, int batch, void *in, double *out);

int cic_paint_f64(hipStream_t s, int nx, int ny, int nz, const double *pos, int64_t n, const double *weight,
                  double wscalar, double *mesh, uint32_t flags) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(n >= 0 && n <= kMaxCells, "particle count %lld out of range", (long long)n);
  JPM_REQUIRE(mesh != nullptr, "mesh is null");
  JPM_REQUIRE(!(flags & JPM_DETERMINISTIC),
              "the float64 paint sums with float64 atomics; JPM_DETERMINISTIC is a float32 mode");
  const bool rel = flags & JPM_RELATIVE;
  const int64_t m = cells(nx, ny, nz);
  JPM_REQUIRE(!rel || n == m, "relative painting needs one particle per lattice node (n=%lld, mesh %lld)",
              (long long)n, (long long)m);
  if (flags & JPM_ZERO_MESH) {
    hipLaunchKernelGGL(zero_f64_kernel, dim3(grid_for(m, 256)), dim3(256), 0, s, mesh, m);
    JPM_LAUNCHED("zero_f64_kernel");
  }
  if (n == 0) return JPM_OK;
  JPM_REQUIRE(pos != nullptr, "positions are null");
  const GridD g = make_grid_d(nx, ny, nz);
  const unsigned grid = grid_for(n, 256, 256 * 128);
  if (rel)
    hipLaunchKernelGGL(paint_f64_kernel<true>, dim3(grid), dim3(256), 0, s, pos, weight, wscalar, uint32_t(n), g, mesh);
  else
    hipLaunchKernelGGL(paint_f64_kernel<false>, dim3(grid), dim3(256), 0, s, pos, weight, wscalar, uint32_t(n), g,
                       mesh);
  JPM_LAUNCHED("paint_f64_kernel");
  return JPM_OK;
}

int cic_read_f64(hipStream_t s, int nx, int ny, int nz, const double *mesh, int ncomp, int64_t comp_stride,
                 const double *pos, int64_t n, double *out, uint32_t flags) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(ncomp >= 1 && ncomp <= 4, "ncomp=%d must be 1..4", ncomp);
  JPM_REQUIRE(n >= 0 && n <= kMaxCells, "particle count %lld out of range", (long long)n);
  const int64_t m = cells(nx, ny, nz);
  JPM_REQUIRE(ncomp == 1 || comp_stride >= m, "component stride %lld smaller than the mesh", (long long)comp_stride);
  const bool rel = flags & JPM_RELATIVE;
  JPM_REQUIRE(!rel || n == m, "relative reading needs one particle per lattice node (n=%lld, mesh %lld)",
              (long long)n, (long long)m);
  if (n == 0) return JPM_OK;
  JPM_REQUIRE(mesh && pos && out, "null pointer argument");
  const GridD g = make_grid_d(nx, ny, nz);
  const unsigned grid = grid_for(n, 256, 256 * 128);
  if (rel)
    hipLaunchKernelGGL(read_f64_kernel<true>, dim3(grid), dim3(256), 0, s, mesh, ncomp, comp_stride, pos, uint32_t(n),
                       g, out);
  else
    hipLaunchKernelGGL(read_f64_kernel<false>, dim3(grid), dim3(256), 0, s, mesh, ncomp, comp_stride, pos,
                       uint32_t(n), g, out);
  JPM_LAUNCHED("read_f64_kernel");
  return JPM_OK;
}

// Workspace of pm_forces_f64: the density mesh, its half spectrum, the three
// force spectra and the three force meshes.
size_t pm_forces_f64_workspace_size(int nx, int ny, int nz) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
  const size_t M = size_t(cells(nx, ny, nz)), H = size_t(half_cells(nx, ny, nz));
  return align_up(M * 8) + align_up(H * 16) + align_up(3 * H * 16) + align_up(3 * M * 8);
}

int pm_forces_f64(hipStream_t s, int nx, int ny, int nz, const double *pos, int64_t n, const double *delta,
                  double r_split, uint32_t flags, uint32_t kflags, double *out3, void *ws, size_t ws_bytes) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(n >= 0 && n <= kMaxCells, "particle count %lld out of range", (long long)n);
  JPM_REQUIRE(out3 != nullptr && (n == 0 || pos != nullptr), "null pointer argument");
  const size_t need = pm_forces_f64_workspace_size(nx, ny, nz);
  JPM_REQUIRE(ws && ws_bytes >= need, "pm_forces_f64 needs %zu workspace bytes, got %zu", need, ws_bytes);
  const int64_t M = cells(nx, ny, nz), H = half_cells(nx, ny, nz);
  char *base = static_cast<char *>(ws);
  double *rho = reinterpret_cast<double *>(base);
  double2 *spec = reinterpret_cast<double2 *>(base + align_up(size_t(M) * 8));
  double2 *spec3 = reinterpret_cast<double2 *>(base + align_up(size_t(M) * 8) + align_up(size_t(H) * 16));
  double *phi3 = reinterpret_cast<double *>(base + align_up(size_t(M) * 8) + align_up(size_t(H) * 16) +
                                            align_up(3 * size_t(H) * 16));
  const double *field = delta;
  if (!field) {
    rc = cic_paint_f64(s, nx, ny, nz, pos, n, nullptr, 1.0, rho, (flags & JPM_RELATIVE) | JPM_ZERO_MESH);
    if (rc) return rc;
    field = rho;
  }
  if ((rc = fft_d2z(s, nx, ny, nz, 1, field, spec))) return rc;
  hipLaunchKernelGGL(force_f64_kernel, dim3(grid_for(H, 256)), dim3(256), 0, s, spec, nx, ny, nz, r_split * r_split,
                     kflags, spec3);
  JPM_LAUNCHED("force_f64_kernel");
  if ((rc = fft_z2d(s, nx, ny, nz, 3, spec3, phi3))) return rc;
  hipLaunchKernelGGL(scale_f64_kernel, dim3(grid_for(3 * M, 256)), dim3(256), 0, s, phi3, 3 * M, 1.0 / double(M));
  JPM_LAUNCHED("scale_f64_kernel");
  return cic_read_f64(s, nx, ny, nz, phi3, 3, M, pos, n, out3, flags & JPM_RELATIVE);
}

}  // namespace jpm
