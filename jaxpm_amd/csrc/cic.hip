// CIC paint and read kernels for gfx950.
//
// Corner maths follows the reference exactly in float32:
//   absolute (painting.py:21-37): f = floor(x); corner c = f + o; weight
//     1 - |x - c| per axis, product over axes; index int32(c) mod n.
//   relative (painting_utils.py:43-69): p = lattice + disp; q = (p + o) mod L
//     (jnp.remainder: fmod, +L when negative); q = floor(q); nd = p - q;
//     nd -= rint(nd / L) * L; weight 1 - |nd|; index q, dropped unless
//     0 <= q < L (scatter drops, gather mode='drop' reads 0).
// The library is built with -ffp-contract=off so these sequences round as
// written.
#include "tiles.hpp"

namespace jpm {

struct Axis {
  int i0, i1;    // wrapped indices of the low / high corner
  float w0, w1;  // CIC weights 1 - |x - c|
  float d0, d1;  // their derivatives d w / d x (|u|' = +1 at u = 0, as jnp.abs)
};

__device__ __forceinline__ int wrap_index(int c, int n) {
  if (uint32_t(c) < uint32_t(n)) return c;   // inside the mesh: no division
  int r = c % n;
  return r < 0 ? r + n : r;
}

__device__ __forceinline__ Axis axis_abs(float x, int n) {
  Axis a;
  float f = floorf(x);
  float f1 = f + 1.0f;
  const float u0 = x - f, u1 = x - f1;
  a.w0 = 1.0f - fabsf(u0);
  a.w1 = 1.0f - fabsf(u1);
  a.d0 = u0 >= 0.0f ? -1.0f : 1.0f;
  a.d1 = u1 >= 0.0f ? -1.0f : 1.0f;
  a.i0 = wrap_index(cell_int(f), n);
  a.i1 = wrap_index(cell_int(f1), n);
  return a;
}

__device__ __forceinline__ float jnp_mod(float x, float L) {
  float m = fmodf(x, L);
  if (m != 0.0f && m < 0.0f) m = m + L;
  return m;
}

__device__ __forceinline__ void corner_rel(float p, float o, float L, int &idx, float &w, float &dw) {
  float q = jnp_mod(p + o, L);
  q = floorf(q);
  float nd = p - q;
  nd = nd - rintf(nd / L) * L;
  // Out-of-range corners (q == L after rounding, or NaN) carry zero weight at
  // a valid index, which is what dropping them does to a sum.
  bool ok = (q >= 0.0f) && (q < L);
  idx = ok ? int(q) : 0;
  w = ok ? 1.0f - fabsf(nd) : 0.0f;
  dw = ok ? (nd >= 0.0f ? -1.0f : 1.0f) : 0.0f;
}

__device__ __forceinline__ Axis axis_rel(float p, float L) {
  Axis a;
  corner_rel(p, 0.0f, L, a.i0, a.w0, a.d0);
  corner_rel(p, 1.0f, L, a.i1, a.w1, a.d1);
  return a;
}

struct Cic {
  uint32_t idx[8];
  float w[8];
};

// Corner order painting.py:24-25: 000,100,010,001,110,101,011,111
// (x is the first axis).
__device__ __forceinline__ void corners_abs_order(const Axis &ax, const Axis &ay, const Axis &az,
                                                  int ny, int nz, Cic &c) {
  const int xs[8] = {0, 1, 0, 0, 1, 1, 0, 1};
  const int ys[8] = {0, 0, 1, 0, 1, 0, 1, 1};
  const int zs[8] = {0, 0, 0, 1, 0, 1, 1, 1};
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    int ix = xs[o] ? ax.i1 : ax.i0;
    int iy = ys[o] ? ay.i1 : ay.i0;
    int iz = zs[o] ? az.i1 : az.i0;
    float wx = xs[o] ? ax.w1 : ax.w0;
    float wy = ys[o] ? ay.w1 : ay.w0;
    float wz = zs[o] ? az.w1 : az.w0;
    c.idx[o] = (uint32_t(ix) * uint32_t(ny) + uint32_t(iy)) * uint32_t(nz) + uint32_t(iz);
    c.w[o] = wx * wy * wz;
  }
}

// Corner order painting_utils.py:43-45: bit d of the corner number is the
// offset along axis d.
__device__ __forceinline__ void corners_bit_order(const Axis &ax, const Axis &ay, const Axis &az,
                                                  int ny, int nz, Cic &c) {
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    int bx = o & 1, by = (o >> 1) & 1, bz = (o >> 2) & 1;
    int ix = bx ? ax.i1 : ax.i0;
    int iy = by ? ay.i1 : ay.i0;
    int iz = bz ? az.i1 : az.i0;
    float wx = bx ? ax.w1 : ax.w0;
    float wy = by ? ay.w1 : ay.w0;
    float wz = bz ? az.w1 : az.w0;
    c.idx[o] = (uint32_t(ix) * uint32_t(ny) + uint32_t(iy)) * uint32_t(nz) + uint32_t(iz);
    c.w[o] = wx * wy * wz;
  }
}

struct Grid {
  int nx, ny, nz;
  FastDiv div_nz, div_ny;
};

template <bool REL>
__device__ __forceinline__ void particle_axes(const float *__restrict__ pos, uint32_t p, const Grid &g,
                                              Axis &ax, Axis &ay, Axis &az) {
  float x = pos[3 * size_t(p) + 0];
  float y = pos[3 * size_t(p) + 1];
  float z = pos[3 * size_t(p) + 2];
  if (REL) {
    uint32_t t = g.div_nz.div(p);
    uint32_t k = p - t * uint32_t(g.nz);
    uint32_t i = g.div_ny.div(t);
    uint32_t j = t - i * uint32_t(g.ny);
    // enmesh: base_indices * cell_size + displacements (float32 add).
    ax = axis_rel(float(i) + x, float(g.nx));
    ay = axis_rel(float(j) + y, float(g.ny));
    az = axis_rel(float(k) + z, float(g.nz));
  } else {
    ax = axis_abs(x, g.nx);
    ay = axis_abs(y, g.ny);
    az = axis_abs(z, g.nz);
  }
}

template <bool REL>
__device__ __forceinline__ void particle_cic(const float *__restrict__ pos, uint32_t p,
                                             const Grid &g, Cic &c) {
  Axis ax, ay, az;
  particle_axes<REL>(pos, p, g, ax, ay, az);
  if (REL)
    corners_bit_order(ax, ay, az, g.ny, g.nz, c);
  else
    corners_abs_order(ax, ay, az, g.ny, g.nz, c);
}

// ------------------------------------------------------------------ paint

template <bool REL, bool HAS_W>
__global__ void __launch_bounds__(256)
paint_atomic_kernel(const float *__restrict__ pos, const float *__restrict__ weight, float wscalar,
                    uint32_t n, Grid g, float *__restrict__ mesh) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    Cic c;
    particle_cic<REL>(pos, p, g, c);
    float wp = HAS_W ? weight[p] : wscalar;
#pragma unroll
    for (int o = 0; o < 8; ++o) unsafeAtomicAdd(&mesh[c.idx[o]], wp * c.w[o]);
  }
}

// Deterministic paint: every deposit is rounded to a 64-bit fixed-point
// integer with a power-of-two scale and summed with integer atomics, which
// are associative, so the result does not depend on the order in which
// particles land.  The scale 2^e keeps sum |w| * 2^e below 2^62 (w over the
// finite weights).  A deposit that is NaN or +-Inf (a non-finite weight or
// position) skips the fixed-point sum and sets *nonfinite; nonfinite_add_kernel
// then adds it to the converted mesh as a float.
template <bool REL, bool HAS_W>
__global__ void __launch_bounds__(256)
paint_fixed_kernel(const float *__restrict__ pos, const float *__restrict__ weight, float wscalar,
                   uint32_t n, Grid g, const float *__restrict__ wmax_dev, float wmax_host,
                   unsigned long long *__restrict__ acc, uint32_t *__restrict__ nonfinite) {
  float wmax = HAS_W ? *wmax_dev : wmax_host;
  double scale = fixed_point_scale(wmax, n);
  bool hit = false;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    Cic c;
    particle_cic<REL>(pos, p, g, c);
    float wp = HAS_W ? weight[p] : wscalar;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const float d = wp * c.w[o];
      if (!isfinite(d)) {
        hit = true;
        continue;
      }
      atomicAdd(&acc[c.idx[o]], fixed_point_q(d, scale));
    }
  }
  if (hit) atomicOr(nonfinite, 1u);
}

// The non-finite deposits of a deterministic paint, added as floats to the
// converted mesh (nc channels, channel c of particle p weighted by
// weight[p * nc + c], or wscalar with no weights).  Exits at once unless the
// paint flagged one.
template <bool REL>
__global__ void __launch_bounds__(256)
nonfinite_add_kernel(const float *__restrict__ pos, const float *__restrict__ weight, float wscalar, int nc,
                     uint32_t n, Grid g, uint32_t m, const uint32_t *__restrict__ nonfinite, float *__restrict__ mesh) {
  if (*nonfinite == 0u) return;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    Cic c;
    particle_cic<REL>(pos, p, g, c);
    for (int ch = 0; ch < nc; ++ch) {
      const float wp = weight ? weight[size_t(p) * nc + ch] : wscalar;
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        const float d = wp * c.w[o];
        if (!isfinite(d)) unsafeAtomicAdd(&mesh[size_t(ch) * m + c.idx[o]], d);
      }
    }
  }
}

__global__ void __launch_bounds__(256)
abs_max_kernel(const float *__restrict__ w, uint32_t n, float *__restrict__ out) {
  __shared__ float sh[4];
  float m = 0.0f;
  // 16-byte loads over the aligned body, scalar loads for the rest
  const uint32_t head = uint32_t((16u - (reinterpret_cast<uintptr_t>(w) & 15u)) & 15u) / 4u;
  const uint32_t h = head < n ? head : n;
  const uint32_t nv = (n - h) / 4u;
  const float4 *w4 = reinterpret_cast<const float4 *>(w + h);
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  for (uint32_t p = tid; p < nv; p += stride) {
    const float4 q = w4[p];
    m = fmaxf(m, fmaxf(fmaxf(finite_abs(q.x), finite_abs(q.y)), fmaxf(finite_abs(q.z), finite_abs(q.w))));
  }
  for (uint32_t p = tid; p < h; p += stride) m = fmaxf(m, finite_abs(w[p]));
  for (uint32_t p = h + 4u * nv + tid; p < n; p += stride) m = fmaxf(m, finite_abs(w[p]));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  // one atomic per workgroup; values are >= 0, so the unsigned bit pattern
  // orders like the float
  if (threadIdx.x == 0) {
    m = fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
    atomicMax(reinterpret_cast<unsigned int *>(out), __float_as_uint(m));
  }
}

__global__ void __launch_bounds__(256)
fixed_to_float_kernel(const unsigned long long *__restrict__ acc, uint32_t m, uint32_t n,
                      const float *__restrict__ wmax_dev, float wmax_host, int has_w,
                      int accumulate, float *__restrict__ mesh) {
  float wmax = has_w ? *wmax_dev : wmax_host;
  double inv = 1.0 / fixed_point_scale(wmax, n);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    float v = float(double((long long)acc[i]) * inv);
    mesh[i] = accumulate ? mesh[i] + v : v;
  }
}

__global__ void __launch_bounds__(256) zero_u64_kernel(unsigned long long *__restrict__ p, uint32_t m) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

__global__ void __launch_bounds__(256) zero_f32_kernel(float *__restrict__ p, uint32_t m) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) p[i] = 0.0f;
}

// ------------------------------------------------------------------- read

template <bool REL, int NC>
__global__ void __launch_bounds__(256)
read_kernel(const float *__restrict__ mesh, uint32_t cstride, const float *__restrict__ pos,
            uint32_t n, Grid g, float *__restrict__ out) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    Cic c;
    particle_cic<REL>(pos, p, g, c);
#pragma unroll
    for (int comp = 0; comp < NC; ++comp) {
      const float *m = mesh + size_t(comp) * cstride;
      float s = 0.0f;
#pragma unroll
      for (int o = 0; o < 8; ++o) s += m[c.idx[o]] * c.w[o];
      out[size_t(p) * NC + comp] = s;
    }
  }
}

// Multi-channel paint: channel c of particle p deposits weight[p*nc + c]
// into mesh + c*M.  This is the transpose of the nc-component read.
template <bool REL, int NC>
__global__ void __launch_bounds__(256)
paint_multi_atomic_kernel(const float *__restrict__ pos, const float *__restrict__ weight, uint32_t n, Grid g,
                          uint32_t m, float *__restrict__ mesh) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    Cic c;
    particle_cic<REL>(pos, p, g, c);
#pragma unroll
    for (int ch = 0; ch < NC; ++ch) {
      const float wp = weight[size_t(p) * NC + ch];
      float *mc = mesh + size_t(ch) * m;
#pragma unroll
      for (int o = 0; o < 8; ++o) unsafeAtomicAdd(&mc[c.idx[o]], wp * c.w[o]);
    }
  }
}

template <bool REL, int NC>
__global__ void __launch_bounds__(256)
paint_multi_fixed_kernel(const float *__restrict__ pos, const float *__restrict__ weight, uint32_t n, Grid g,
                         uint32_t m, const float *__restrict__ wmax_dev, unsigned long long *__restrict__ acc,
                         uint32_t *__restrict__ nonfinite) {
  const double scale = fixed_point_scale(*wmax_dev, n);
  bool hit = false;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    Cic c;
    particle_cic<REL>(pos, p, g, c);
#pragma unroll
    for (int ch = 0; ch < NC; ++ch) {
      const float wp = weight[size_t(p) * NC + ch];
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        const float d = wp * c.w[o];
        if (!isfinite(d)) {
          hit = true;
          continue;
        }
        atomicAdd(&acc[size_t(ch) * m + c.idx[o]], fixed_point_q(d, scale));
      }
    }
  }
  if (hit) atomicOr(nonfinite, 1u);
}

// Gradient of CIC sampling with respect to the particle coordinate:
//   gpos[p, e] (+)= w_p * sum_c cot[p, c] * sum_o mesh_c[idx_o] dW_o/dx_e
// cot == NULL means cot = 1.  This is the position cotangent of both
// cic_read (mesh = the sampled field, cot = output cotangent) and cic_paint
// (mesh = cotangent of the painted mesh, w = particle weight).
template <bool REL, int NC>
__global__ void __launch_bounds__(256)
grad_pos_kernel(const float *__restrict__ mesh, uint32_t cstride, const float *__restrict__ pos, uint32_t n,
                Grid g, const float *__restrict__ cot, const float *__restrict__ weight, float wscalar,
                int accumulate, float *__restrict__ gpos) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    Axis ax, ay, az;
    particle_axes<REL>(pos, p, g, ax, ay, az);
    const uint32_t ixs[2] = {uint32_t(ax.i0), uint32_t(ax.i1)};
    const uint32_t iys[2] = {uint32_t(ay.i0), uint32_t(ay.i1)};
    const uint32_t izs[2] = {uint32_t(az.i0), uint32_t(az.i1)};
    const float wx[2] = {ax.w0, ax.w1}, wy[2] = {ay.w0, ay.w1}, wz[2] = {az.w0, az.w1};
    const float dx[2] = {ax.d0, ax.d1}, dy[2] = {ay.d0, ay.d1}, dz[2] = {az.d0, az.d1};
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll
    for (int comp = 0; comp < NC; ++comp) {
      const float *mc = mesh + size_t(comp) * cstride;
      float sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        const int bx = o & 1, by = (o >> 1) & 1, bz = (o >> 2) & 1;
        const float v = mc[(ixs[bx] * uint32_t(g.ny) + iys[by]) * uint32_t(g.nz) + izs[bz]];
        sx += v * (dx[bx] * wy[by] * wz[bz]);
        sy += v * (wx[bx] * dy[by] * wz[bz]);
        sz += v * (wx[bx] * wy[by] * dz[bz]);
      }
      const float cc = cot ? cot[size_t(p) * NC + comp] : 1.0f;
      gx += cc * sx;
      gy += cc * sy;
      gz += cc * sz;
    }
    const float wp = weight ? weight[p] : wscalar;
    gx *= wp;
    gy *= wp;
    gz *= wp;
    float *o = gpos + 3 * size_t(p);
    if (accumulate) {
      gx += o[0];
      gy += o[1];
      gz += o[2];
    }
    o[0] = gx;
    o[1] = gy;
    o[2] = gz;
  }
}

// Position cotangent of the batched paint: gpos_p = sum_o g[idx_o] dW_o/dx
// over each simulation's particles and mesh (blockIdx.y: simulation).
template <bool REL>
__global__ void __launch_bounds__(256)
grad_pos_batch_kernel(const SimRef *__restrict__ sims, const float *__restrict__ mesh_all,
                      const float *__restrict__ pos_all, float *__restrict__ gpos_all) {
  const SimRef &s = sims[blockIdx.y];
  Grid g;
  g.nx = s.t.nx;
  g.ny = s.t.ny;
  g.nz = s.t.nz;
  g.div_nz = s.t.div_nz;
  g.div_ny = s.t.div_ny;
  const float *mesh = mesh_all + s.cell0;
  const float *pos = pos_all + 3 * size_t(s.part0);
  float *gpos = gpos_all + 3 * size_t(s.part0);
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < s.n; p += gridDim.x * blockDim.x) {
    Axis ax, ay, az;
    particle_axes<REL>(pos, p, g, ax, ay, az);
    const uint32_t ixs[2] = {uint32_t(ax.i0), uint32_t(ax.i1)};
    const uint32_t iys[2] = {uint32_t(ay.i0), uint32_t(ay.i1)};
    const uint32_t izs[2] = {uint32_t(az.i0), uint32_t(az.i1)};
    const float wx[2] = {ax.w0, ax.w1}, wy[2] = {ay.w0, ay.w1}, wz[2] = {az.w0, az.w1};
    const float dx[2] = {ax.d0, ax.d1}, dy[2] = {ay.d0, ay.d1}, dz[2] = {az.d0, az.d1};
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int bx = o & 1, by = (o >> 1) & 1, bz = (o >> 2) & 1;
      const float v = mesh[(ixs[bx] * uint32_t(g.ny) + iys[by]) * uint32_t(g.nz) + izs[bz]];
      sx += v * (dx[bx] * wy[by] * wz[bz]);
      sy += v * (wx[bx] * dy[by] * wz[bz]);
      sz += v * (wx[bx] * wy[by] * dz[bz]);
    }
    float *o = gpos + 3 * size_t(p);
    o[0] = 1.0f * sx;
    o[1] = 1.0f * sy;
    o[2] = 1.0f * sz;
  }
}

int grad_pos_batch(hipStream_t s, const SimRef *sims, int B, bool rel, const float *mesh, const float *pos,
                   float *gpos) {
  JPM_REQUIRE(sims && mesh && pos && gpos && B >= 1, "null pointer argument");
  JPM_TIMED(s, "grad_pos_kernel");
  if (rel)
    hipLaunchKernelGGL(grad_pos_batch_kernel<true>, dim3(256, B), dim3(256), 0, s, sims, mesh, pos, gpos);
  else
    hipLaunchKernelGGL(grad_pos_batch_kernel<false>, dim3(256, B), dim3(256), 0, s, sims, mesh, pos, gpos);
  JPM_LAUNCHED("grad_pos_batch_kernel");
  return JPM_OK;
}

// ------------------------------------------------------------ host side

static Grid make_grid(int nx, int ny, int nz) {
  Grid g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.div_nz = FastDiv(uint32_t(nz));
  g.div_ny = FastDiv(uint32_t(ny));
  return g;
}

// *out = max |w[i]| over n floats (out is zeroed first).
int abs_max(hipStream_t s, const float *w, uint32_t n, float *out) {
  hipLaunchKernelGGL(zero_f32_kernel, dim3(1), dim3(64), 0, s, out, 1u);
  {
    JPM_TIMED(s, "abs_max_kernel");
    hipLaunchKernelGGL(abs_max_kernel, dim3(grid_for((int64_t(n) + 3) / 4, 256, 2048)), dim3(256), 0, s, w, n, out);
  }
  JPM_LAUNCHED("abs_max_kernel");
  return JPM_OK;
}

int zero_f32(hipStream_t s, float *p, int64_t m) {
  hipLaunchKernelGGL(zero_f32_kernel, dim3(grid_for(m, 256)), dim3(256), 0, s, p, uint32_t(m));
  JPM_LAUNCHED("zero_f32_kernel");
  return JPM_OK;
}

int cic_paint(hipStream_t s, int nx, int ny, int nz, const float *pos, int64_t n, const float *weight,
              float wscalar, float *mesh, uint32_t flags, void *ws, size_t ws_bytes) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(n >= 0 && n <= kMaxCells, "particle count %lld out of range", (long long)n);
  JPM_REQUIRE(mesh != nullptr, "mesh is null");
  JPM_REQUIRE(n == 0 || pos != nullptr, "positions are null");
  const bool rel = flags & JPM_RELATIVE;
  const int64_t m = cells(nx, ny, nz);
  JPM_REQUIRE(!rel || n == m, "relative painting needs one particle per lattice node (n=%lld, mesh %lld)",
              (long long)n, (long long)m);
  if (flags & JPM_ZERO_MESH) {
    rc = zero_f32(s, mesh, m);
    if (rc) return rc;
  }
  if (n == 0) return JPM_OK;
  Grid g = make_grid(nx, ny, nz);
  const unsigned grid = grid_for(n, 256, 256 * 128);
  const bool hw = weight != nullptr;
  if (!(flags & JPM_DETERMINISTIC)) {
#define JPM_PAINT(R, W)                                                                          \
  hipLaunchKernelGGL((paint_atomic_kernel<R, W>), dim3(grid), dim3(256), 0, s, pos, weight, wscalar, \
                     uint32_t(n), g, mesh)
    if (rel) {
      if (hw) JPM_PAINT(true, true); else JPM_PAINT(true, false);
    } else {
      if (hw) JPM_PAINT(false, true); else JPM_PAINT(false, false);
    }
#undef JPM_PAINT
    JPM_LAUNCHED("paint_atomic_kernel");
    return JPM_OK;
  }
  const size_t need = jpm_cic_paint_workspace_size(nx, ny, nz, n, flags);
  JPM_REQUIRE(ws != nullptr && ws_bytes >= need, "deterministic paint needs %zu workspace bytes, got %zu",
              need, ws_bytes);
  auto *acc = static_cast<unsigned long long *>(ws);
  // [0] max |finite weight|, [1] the non-finite flag
  float *wmax = reinterpret_cast<float *>(static_cast<char *>(ws) + align_up(size_t(m) * 8));
  uint32_t *nonfinite = reinterpret_cast<uint32_t *>(wmax + 1);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_for(m, 256)), dim3(256), 0, s, acc, uint32_t(m));
  JPM_LAUNCHED("zero_u64_kernel");
  hipLaunchKernelGGL(zero_f32_kernel, dim3(1), dim3(64), 0, s, wmax, 2u);
  if (hw) {
    {
      JPM_TIMED(s, "abs_max_kernel");
      hipLaunchKernelGGL(abs_max_kernel, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, weight, uint32_t(n), wmax);
    }
    JPM_LAUNCHED("abs_max_kernel");
  }
  const float wh = finite_abs(wscalar);
#define JPM_PAINT(R, W)                                                                          \
  hipLaunchKernelGGL((paint_fixed_kernel<R, W>), dim3(grid), dim3(256), 0, s, pos, weight, wscalar,  \
                     uint32_t(n), g, wmax, wh, acc, nonfinite)
  if (rel) {
    if (hw) JPM_PAINT(true, true); else JPM_PAINT(true, false);
  } else {
    if (hw) JPM_PAINT(false, true); else JPM_PAINT(false, false);
  }
#undef JPM_PAINT
  JPM_LAUNCHED("paint_fixed_kernel");
  hipLaunchKernelGGL(fixed_to_float_kernel, dim3(grid_for(m, 256)), dim3(256), 0, s, acc, uint32_t(m),
                     uint32_t(n), wmax, wh, int(hw), int(!(flags & JPM_ZERO_MESH)), mesh);
  JPM_LAUNCHED("fixed_to_float_kernel");
  if (rel)
    hipLaunchKernelGGL(nonfinite_add_kernel<true>, dim3(grid), dim3(256), 0, s, pos, weight, wscalar, 1,
                       uint32_t(n), g, uint32_t(m), nonfinite, mesh);
  else
    hipLaunchKernelGGL(nonfinite_add_kernel<false>, dim3(grid), dim3(256), 0, s, pos, weight, wscalar, 1,
                       uint32_t(n), g, uint32_t(m), nonfinite, mesh);
  JPM_LAUNCHED("nonfinite_add_kernel");
  return JPM_OK;
}

int cic_read(hipStream_t s, int nx, int ny, int nz, const float *mesh, int ncomp, int64_t comp_stride,
             const float *pos, int64_t n, float *out, uint32_t flags) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(ncomp >= 1 && ncomp <= 4, "ncomp=%d must be 1..4", ncomp);
  JPM_REQUIRE(n >= 0 && n <= kMaxCells, "particle count %lld out of range", (long long)n);
  const int64_t m = cells(nx, ny, nz);
  JPM_REQUIRE(ncomp == 1 || (comp_stride >= m && comp_stride <= kMaxCells * 2),
              "component stride %lld smaller than the mesh", (long long)comp_stride);
  JPM_REQUIRE(int64_t(ncomp - 1) * comp_stride + m <= (int64_t(1) << 32),
              "components exceed the 32-bit offset range");
  const bool rel = flags & JPM_RELATIVE;
  JPM_REQUIRE(!rel || n == m, "relative reading needs one particle per lattice node (n=%lld, mesh %lld)",
              (long long)n, (long long)m);
  if (n == 0) return JPM_OK;
  JPM_REQUIRE(mesh && pos && out, "null pointer argument");
  Grid g = make_grid(nx, ny, nz);
  const unsigned grid = grid_for(n, 256, 256 * 128);
  const uint32_t cs = uint32_t(comp_stride);
#define JPM_READ(R, C)                                                                      \
  hipLaunchKernelGGL((read_kernel<R, C>), dim3(grid), dim3(256), 0, s, mesh, cs, pos, uint32_t(n), g, \
                     out)
#define JPM_READ_NC(R)                 \
  switch (ncomp) {                     \
    case 1: JPM_READ(R, 1); break;     \
    case 2: JPM_READ(R, 2); break;     \
    case 3: JPM_READ(R, 3); break;     \
    default: JPM_READ(R, 4); break;    \
  }
  if (rel) {
    JPM_READ_NC(true);
  } else {
    JPM_READ_NC(false);
  }
#undef JPM_READ_NC
#undef JPM_READ
  JPM_LAUNCHED("read_kernel");
  return JPM_OK;
}

int cic_paint_multi(hipStream_t s, int nx, int ny, int nz, const float *pos, int64_t n, const float *weight,
                    int nc, float *mesh, uint32_t flags, void *ws, size_t ws_bytes) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(nc >= 1 && nc <= 4, "channel count %d must be 1..4", nc);
  JPM_REQUIRE(n >= 0 && n <= kMaxCells, "particle count %lld out of range", (long long)n);
  const int64_t m = cells(nx, ny, nz);
  JPM_REQUIRE(int64_t(nc) * m <= (int64_t(1) << 32), "channels exceed the 32-bit offset range");
  const bool rel = flags & JPM_RELATIVE;
  JPM_REQUIRE(!rel || n == m, "relative painting needs one particle per lattice node (n=%lld, mesh %lld)",
              (long long)n, (long long)m);
  JPM_REQUIRE(mesh != nullptr, "mesh is null");
  if (flags & JPM_ZERO_MESH) {
    rc = zero_f32(s, mesh, nc * m);
    if (rc) return rc;
  }
  if (n == 0) return JPM_OK;
  JPM_REQUIRE(pos && weight, "null pointer argument");
  Grid g = make_grid(nx, ny, nz);
  const unsigned grid = grid_for(n, 256, 256 * 128);
  if (!(flags & JPM_DETERMINISTIC)) {
#define JPM_PM(R, C)                                                                                   \
  hipLaunchKernelGGL((paint_multi_atomic_kernel<R, C>), dim3(grid), dim3(256), 0, s, pos, weight, uint32_t(n), \
                     g, uint32_t(m), mesh)
#define JPM_PM_NC(R)              \
  switch (nc) {                   \
    case 1: JPM_PM(R, 1); break;  \
    case 2: JPM_PM(R, 2); break;  \
    case 3: JPM_PM(R, 3); break;  \
    default: JPM_PM(R, 4); break; \
  }
    if (rel) {
      JPM_PM_NC(true);
    } else {
      JPM_PM_NC(false);
    }
#undef JPM_PM_NC
#undef JPM_PM
    JPM_LAUNCHED("paint_multi_atomic_kernel");
    return JPM_OK;
  }
  const size_t need = align_up(size_t(nc) * size_t(m) * 8) + 256;
  JPM_REQUIRE(ws != nullptr && ws_bytes >= need, "deterministic paint needs %zu workspace bytes, got %zu", need,
              ws_bytes);
  auto *acc = static_cast<unsigned long long *>(ws);
  // [0] max |finite weight|, [1] the non-finite flag
  float *wmax = reinterpret_cast<float *>(static_cast<char *>(ws) + align_up(size_t(nc) * size_t(m) * 8));
  uint32_t *nonfinite = reinterpret_cast<uint32_t *>(wmax + 1);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_for(nc * m, 256)), dim3(256), 0, s, acc, uint32_t(nc * m));
  hipLaunchKernelGGL(zero_f32_kernel, dim3(1), dim3(64), 0, s, wmax, 2u);
  {
    JPM_TIMED(s, "abs_max_kernel");
    hipLaunchKernelGGL(abs_max_kernel, dim3(grid_for(n * nc, 256, 2048)), dim3(256), 0, s, weight,
                       uint32_t(n * nc), wmax);
  }
  JPM_LAUNCHED("abs_max_kernel");
#define JPM_PM(R, C)                                                                                  \
  hipLaunchKernelGGL((paint_multi_fixed_kernel<R, C>), dim3(grid), dim3(256), 0, s, pos, weight, uint32_t(n), \
                     g, uint32_t(m), wmax, acc, nonfinite)
#define JPM_PM_NC(R)              \
  switch (nc) {                   \
    case 1: JPM_PM(R, 1); break;  \
    case 2: JPM_PM(R, 2); break;  \
    case 3: JPM_PM(R, 3); break;  \
    default: JPM_PM(R, 4); break; \
  }
  if (rel) {
    JPM_PM_NC(true);
  } else {
    JPM_PM_NC(false);
  }
#undef JPM_PM_NC
#undef JPM_PM
  JPM_LAUNCHED("paint_multi_fixed_kernel");
  hipLaunchKernelGGL(fixed_to_float_kernel, dim3(grid_for(nc * m, 256)), dim3(256), 0, s, acc, uint32_t(nc * m),
                     uint32_t(n), wmax, 0.0f, 1, int(!(flags & JPM_ZERO_MESH)), mesh);
  JPM_LAUNCHED("fixed_to_float_kernel");
  if (rel)
    hipLaunchKernelGGL(nonfinite_add_kernel<true>, dim3(grid), dim3(256), 0, s, pos, weight, 0.0f, nc, uint32_t(n),
                       g, uint32_t(m), nonfinite, mesh);
  else
    hipLaunchKernelGGL(nonfinite_add_kernel<false>, dim3(grid), dim3(256), 0, s, pos, weight, 0.0f, nc, uint32_t(n),
                       g, uint32_t(m), nonfinite, mesh);
  JPM_LAUNCHED("nonfinite_add_kernel");
  return JPM_OK;
}

int cic_grad_pos(hipStream_t s, int nx, int ny, int nz, const float *mesh, int ncomp, int64_t comp_stride,
                 const float *pos, int64_t n, const float *cot, const float *weight, float wscalar, float *gpos,
                 uint32_t flags, int accumulate) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(ncomp >= 1 && ncomp <= 4, "ncomp=%d must be 1..4", ncomp);
  JPM_REQUIRE(n >= 0 && n <= kMaxCells, "particle count %lld out of range", (long long)n);
  const int64_t m = cells(nx, ny, nz);
  JPM_REQUIRE(ncomp == 1 || comp_stride >= m, "component stride %lld smaller than the mesh", (long long)comp_stride);
  JPM_REQUIRE(int64_t(ncomp - 1) * comp_stride + m <= (int64_t(1) << 32), "components exceed the 32-bit offset range");
  const bool rel = flags & JPM_RELATIVE;
  JPM_REQUIRE(!rel || n == m, "relative mode needs one particle per lattice node (n=%lld, mesh %lld)",
              (long long)n, (long long)m);
  if (n == 0) return JPM_OK;
  JPM_REQUIRE(mesh && pos && gpos, "null pointer argument");
  Grid g = make_grid(nx, ny, nz);
  const unsigned grid = grid_for(n, 256, 256 * 128);
  const uint32_t cs = uint32_t(ncomp == 1 ? m : comp_stride);
#define JPM_G(R, C)                                                                                          \
  hipLaunchKernelGGL((grad_pos_kernel<R, C>), dim3(grid), dim3(256), 0, s, mesh, cs, pos, uint32_t(n), g, cot, \
                     weight, wscalar, accumulate, gpos)
#define JPM_G_NC(R)              \
  switch (ncomp) {               \
    case 1: JPM_G(R, 1); break;  \
    case 2: JPM_G(R, 2); break;  \
    case 3: JPM_G(R, 3); break;  \
    default: JPM_G(R, 4); break; \
  }
  if (rel) {
    JPM_G_NC(true);
  } else {
    JPM_G_NC(false);
  }
#undef JPM_G_NC
#undef JPM_G
  JPM_LAUNCHED("grad_pos_kernel");
  return JPM_OK;
}

}  // namespace jpm

namespace jpm {

// ----------------------------------------------------- integrator kernels

// y += a * x over n floats (drift x += beta v; adjoint lambda_v += beta lambda_x).
__global__ void __launch_bounds__(256) axpy_kernel(float a, const float4 *__restrict__ x, float4 *__restrict__ y,
                                                   uint32_t n4) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += gridDim.x * blockDim.x) {
    float4 u = x[i], v = y[i];
    v.x += a * u.x;
    v.y += a * u.y;
    v.z += a * u.z;
    v.w += a * u.w;
    y[i] = v;
  }
}

__global__ void axpy_tail_kernel(float a, const float *__restrict__ x, float *__restrict__ y, uint32_t start,
                                 uint32_t n) {
  uint32_t i = start + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

// Kick fused into the 3-component read: vel[p] += gamma * sum_o phi[idx_o] W_o.
template <bool REL>
__global__ void __launch_bounds__(256)
kick_kernel(const float *__restrict__ phi, uint32_t cstride, const float *__restrict__ pos, uint32_t n, Grid g,
            float gamma, float *__restrict__ vel) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    Cic c;
    particle_cic<REL>(pos, p, g, c);
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      const float *m = phi + size_t(comp) * cstride;
      float s = 0.0f;
#pragma unroll
      for (int o = 0; o < 8; ++o) s += m[c.idx[o]] * c.w[o];
      vel[size_t(p) * 3 + comp] += gamma * s;
    }
  }
}

// Rows of 3 floats in another order, one or two arrays per launch:
// out[k] = in[order[k]] (gather) or out[order[k]] = in[k] (scatter, the
// inverse).  Entries of order >= n are skipped, so a malformed order cannot
// address outside the arrays.
template <bool SCATTER>
__global__ void __launch_bounds__(256)
permute_rows_kernel(const uint32_t *__restrict__ order, const float *__restrict__ in0, float *__restrict__ out0,
                    const float *__restrict__ in1, float *__restrict__ out1, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t row = order[k];
  if (row >= n) return;
  const size_t src = 3 * size_t(SCATTER ? k : row), dst = 3 * size_t(SCATTER ? row : k);
  const float a0 = in0[src], a1 = in0[src + 1], a2 = in0[src + 2];
  if (in1) {
    const float b0 = in1[src], b1 = in1[src + 1], b2 = in1[src + 2];
    out1[dst] = b0;
    out1[dst + 1] = b1;
    out1[dst + 2] = b2;
  }
  out0[dst] = a0;
  out0[dst + 1] = a1;
  out0[dst + 2] = a2;
}

int permute_rows(hipStream_t s, int64_t n, const uint32_t *order, const float *in0, float *out0, const float *in1,
                 float *out1, int scatter) {
  JPM_REQUIRE(n >= 0 && n <= int64_t(0xffffffffu) - 255, "row count %lld out of range", (long long)n);
  if (n == 0) return JPM_OK;
  JPM_REQUIRE(order && in0 && out0, "null pointer argument");
  JPM_REQUIRE(!in1 == !out1, "the second array needs both its input and its output");
  JPM_REQUIRE(in0 != out0 && (!in1 || in1 != out1), "the permutation is out of place");
  JPM_TIMED(s, "permute_rows_kernel");
  hipLaunchKernelGGL(scatter ? permute_rows_kernel<true> : permute_rows_kernel<false>,
                     dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s, order, in0, out0, in1, out1, uint32_t(n));
  JPM_LAUNCHED("permute_rows_kernel");
  return JPM_OK;
}

int axpy(hipStream_t s, int64_t n, float a, const float *x, float *y) {
  JPM_REQUIRE(n >= 0 && n <= (int64_t(1) << 32) - 1, "length %lld out of range", (long long)n);
  if (n == 0) return JPM_OK;
  JPM_REQUIRE(x && y, "null pointer argument");
  JPM_REQUIRE((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) % 16 == 0,
              "axpy operands must be 16-byte aligned");
  const uint32_t n4 = uint32_t(n / 4);
  if (n4) {
    {
      JPM_TIMED(s, "axpy_kernel");
      hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n4, 256, 256 * 128)), dim3(256), 0, s, a,
                         reinterpret_cast<const float4 *>(x), reinterpret_cast<float4 *>(y), n4);
    }
    JPM_LAUNCHED("axpy_kernel");
  }
  if (uint32_t(n) > 4 * n4) {
    hipLaunchKernelGGL(axpy_tail_kernel, dim3(1), dim3(64), 0, s, a, x, y, 4 * n4, uint32_t(n));
    JPM_LAUNCHED("axpy_tail_kernel");
  }
  return JPM_OK;
}

int kick(hipStream_t s, int nx, int ny, int nz, const float *phi3, const float *pos, int64_t n, float gamma,
         float *vel, uint32_t flags) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(n >= 0 && n <= kMaxCells, "particle count %lld out of range", (long long)n);
  const int64_t m = cells(nx, ny, nz);
  JPM_REQUIRE(3 * m <= (int64_t(1) << 32), "mesh too large");
  const bool rel = flags & JPM_RELATIVE;
  JPM_REQUIRE(!rel || n == m, "relative mode needs one particle per lattice node");
  if (n == 0) return JPM_OK;
  JPM_REQUIRE(phi3 && pos && vel, "null pointer argument");
  Grid g = make_grid(nx, ny, nz);
  const unsigned grid = grid_for(n, 256, 256 * 128);
  if (rel)
    {
      JPM_TIMED(s, "kick_kernel");
      hipLaunchKernelGGL(kick_kernel<true>, dim3(grid), dim3(256), 0, s, phi3, uint32_t(m), pos, uint32_t(n), g, gamma,
                         vel);
    }
  else
    {
      JPM_TIMED(s, "kick_kernel");
      hipLaunchKernelGGL(kick_kernel<false>, dim3(grid), dim3(256), 0, s, phi3, uint32_t(m), pos, uint32_t(n), g,
                         gamma, vel);
    }
  JPM_LAUNCHED("kick_kernel");
  return JPM_OK;
}

}  // namespace jpm
