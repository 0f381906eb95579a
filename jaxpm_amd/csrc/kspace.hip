// Fused Fourier-space passes over hipFFT half spectra ([nx][ny][nz/2+1]
// complex64).  Each pass reads the spectrum once and writes its outputs once;
// per-axis multiplier tables (kernels.py:10-136 evaluated in float64 at the
// float32 wave numbers, rounded to float32) live in a cached device buffer.
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "kspace.hpp"
#include "tiles.hpp"

namespace jpm {

static std::mutex g_tab_mu;
static std::map<std::tuple<int, int, int, uint32_t>, float *> g_tabs;

int get_tabs(hipStream_t st, int nx, int ny, int nz, uint32_t flags, Tabs &t) {
  const int nzh = nz / 2 + 1;
  const size_t len = size_t(nx) + ny + nzh;
  const uint32_t key_flags = flags & (JPM_K_GRAD_EXACT | JPM_K_LAPL_FD);
  float *dev = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto it = g_tabs.find({nx, ny, nz, key_flags});
    if (it != g_tabs.end()) dev = it->second;
  }
  if (!dev) {
    std::vector<float> h(5 * len);
    const int ns[3] = {nx, ny, nz};
    const int lens[3] = {nx, ny, nzh};
    size_t off = 0;
    for (int d = 0; d < 3; ++d) {
      const int n = ns[d];
      for (int i = 0; i < lens[d]; ++i) {
        const int f = (i < (n + 1) / 2) ? i : i - n;  // fftfreq ordering
        const float w32 = float(2.0 * M_PI * double(f) / double(n));
        const double w = double(w32);
        const double sinc = (w == 0.0) ? 1.0 : std::sin(w / 2.0) / (w / 2.0);
        double g;
        if (key_flags & JPM_K_GRAD_EXACT)
          g = (i == n / 2) ? 0.0 : w;  // kernels.py:57-61 zeroes entry len//2
        else
          g = (8.0 * std::sin(w) - std::sin(2.0 * w)) / 6.0;  // kernels.py:63-65
        const double k2 = (key_flags & JPM_K_LAPL_FD) ? (w * sinc) * (w * sinc) : w * w;
        h[0 * len + off + i] = w32;
        h[1 * len + off + i] = float(k2);
        h[2 * len + off + i] = float(w * w);
        h[3 * len + off + i] = float(g);
        h[4 * len + off + i] = float(sinc);
      }
      off += lens[d];
    }
    float *buf = nullptr;
    JPM_HIP(hipMalloc(&buf, h.size() * sizeof(float)));
    hipError_t e = hipMemcpyAsync(buf, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      (void)hipFree(buf);
      return fail(JPM_E_HIP, "uploading wave-number tables failed: %s", hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto ins = g_tabs.emplace(std::make_tuple(nx, ny, nz, key_flags), buf);
    if (!ins.second) {
      (void)hipFree(buf);
      buf = ins.first->second;
    }
    dev = buf;
  }
  t.k = dev;
  t.k2 = dev + len;
  t.k2e = dev + 2 * len;
  t.g = dev + 3 * len;
  t.s = dev + 4 * len;
  t.oy = uint32_t(nx);
  t.oz = uint32_t(nx + ny);
  return JPM_OK;
}

void free_tabs() {
  std::lock_guard<std::mutex> lk(g_tab_mu);
  for (auto &kv : g_tabs) (void)hipFree(kv.second);
  g_tabs.clear();
}

struct HalfGrid {
  uint32_t H, ny, nzh;
  FastDiv div_nzh, div_ny;
  __device__ __forceinline__ void coords(uint32_t e, uint32_t &i, uint32_t &j, uint32_t &k) const {
    uint32_t r = div_nzh.div(e);
    k = e - r * nzh;
    i = div_ny.div(r);
    j = r - i * ny;
  }
};

static HalfGrid make_half(int nx, int ny, int nz) {
  HalfGrid h;
  h.nzh = uint32_t(nz / 2 + 1);
  h.ny = uint32_t(ny);
  h.H = uint32_t(half_cells(nx, ny, nz));
  h.div_nzh = FastDiv(h.nzh);
  h.div_ny = FastDiv(h.ny);
  return h;
}

__global__ void __launch_bounds__(256)
force_kernel(const float2 *__restrict__ dk, float2 *__restrict__ out, HalfGrid hg, Tabs t, float r2,
             float scale, int deconv) {
  // blockIdx.y: spectrum of a batch of equal shapes (in stride H, out 3 H)
  dk += size_t(blockIdx.y) * hg.H;
  out += 3 * size_t(blockIdx.y) * hg.H;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < hg.H; e += gridDim.x * blockDim.x) {
    uint32_t i, j, k;
    hg.coords(e, i, j, k);
    const uint32_t jy = t.oy + j, kz = t.oz + k;
    float kk = (t.k2[i] + t.k2[jy]) + t.k2[kz];
    float f = green(kk);
    if (r2 != 0.0f) {
      float kke = (t.k2e[i] + t.k2e[jy]) + t.k2e[kz];
      f = f * expf(-kke * r2);
    }
    if (deconv) {
      float p = (t.s[i] * t.s[jy]) * t.s[kz];
      f = f / (p * p);
    }
    f = f * scale;
    float2 v = dk[e];
    float pr = v.x * f, pi = v.y * f;
    // -(i g) (pr + i pi) = g pi - i g pr
    float gx = t.g[i], gy = t.g[jy], gz = t.g[kz];
    out[e] = make_float2(gx * pi, -(gx * pr));
    out[size_t(hg.H) + e] = make_float2(gy * pi, -(gy * pr));
    out[2 * size_t(hg.H) + e] = make_float2(gz * pi, -(gz * pr));
  }
}

// Potential only: out = scale * G Lr [C] delta_k (the real-space 4-point
// stencil then takes the gradient).
__global__ void __launch_bounds__(256)
potential_kernel(const float2 *__restrict__ dk, float2 *__restrict__ out, HalfGrid hg, Tabs t, float r2,
                 float scale_s, const float *__restrict__ scale_dev, int deconv) {
  // blockIdx.y: spectrum of a batch of equal shapes (stride H), scaled by
  // scale_dev[y] when given
  dk += size_t(blockIdx.y) * hg.H;
  out += size_t(blockIdx.y) * hg.H;
  const float scale = scale_dev ? scale_dev[blockIdx.y] : scale_s;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < hg.H; e += gridDim.x * blockDim.x) {
    uint32_t i, j, k;
    hg.coords(e, i, j, k);
    const float f = potential_factor(t, i, t.oy + j, t.oz + k, r2, deconv, scale);
    const float2 v = dk[e];
    out[e] = make_float2(v.x * f, v.y * f);
  }
}

__global__ void __launch_bounds__(256)
compensate_kernel(float2 *__restrict__ fk, HalfGrid hg, Tabs t, float scale) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < hg.H; e += gridDim.x * blockDim.x) {
    uint32_t i, j, k;
    hg.coords(e, i, j, k);
    float p = (t.s[i] * t.s[t.oy + j]) * t.s[t.oz + k];
    float c = 1.0f / (p * p) * scale;
    float2 v = fk[e];
    fk[e] = make_float2(v.x * c, v.y * c);
  }
}

// jnp.interp(x, xp, fp) with the default clamping (fp[0] left of xp[0],
// fp[-1] right of xp[-1]).
__device__ __forceinline__ float interp(float x, const float *__restrict__ xp, const float *__restrict__ fp,
                                        int n) {
  // i = clip(searchsorted(xp, x, side='right'), 1, n - 1)
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (xp[mid] <= x) lo = mid + 1; else hi = mid;
  }
  int i = lo < 1 ? 1 : (lo > n - 1 ? n - 1 : lo);
  float df = fp[i] - fp[i - 1];
  float dx = xp[i] - xp[i - 1];
  float delta = x - xp[i - 1];
  float f = (fabsf(dx) <= 0.0f) ? fp[i - 1] : fp[i - 1] + (delta / dx) * df;
  if (x < xp[0]) f = fp[0];
  if (x > xp[n - 1]) f = fp[n - 1];
  return f;
}

__global__ void __launch_bounds__(256)
linear_field_kernel(float2 *__restrict__ fk, HalfGrid hg, Tabs t, float3 box, float3 nm, float mcells, float vol,
                    const float *__restrict__ ktab, const float *__restrict__ pktab, int ntab, float scale,
                    const float *__restrict__ boxes) {
  // blockIdx.y: spectrum of a batch of equal shapes (stride H) with its own
  // box (boxes[3 y..3 y+2], when given) and P(k) table (pktab + y ntab)
  fk += size_t(blockIdx.y) * hg.H;
  pktab += size_t(blockIdx.y) * ntab;
  if (boxes) {
    box = make_float3(boxes[3 * blockIdx.y], boxes[3 * blockIdx.y + 1], boxes[3 * blockIdx.y + 2]);
    vol = float(double(box.x) * double(box.y) * double(box.z));
  }
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < hg.H; e += gridDim.x * blockDim.x) {
    uint32_t i, j, k;
    hg.coords(e, i, j, k);
    // pm.py:137-138: (k_i / L_i * n_i)^2 summed, then ** 0.5
    float ax = t.k[i] / box.x * nm.x, ay = t.k[t.oy + j] / box.y * nm.y, az = t.k[t.oz + k] / box.z * nm.z;
    float kmesh = sqrtf((ax * ax + ay * ay) + az * az);
    float pk = interp(kmesh, ktab, pktab, ntab);
    float c = sqrtf(pk * mcells / vol) * scale;
    float2 v = fk[e];
    fk[e] = make_float2(v.x * c, v.y * c);
  }
}

// Cotangent of the linear field's power-spectrum table.  With c(k) =
// sqrt(pk(|k|) M / V) * scale the forward output is C2R(W c) for the noise
// spectrum W, so for a real-space cotangent with spectrum G,
//   dL/dpk(k) = mult(k) Re(W conj G) c / (2 pk)   (mult 2 for the modes whose
// Hermitian mirror is not stored), and pk(k) = (1 - t) pktab[i-1] + t pktab[i]
// (jnp.interp; end values outside the table).  Lanes of one table interval
// are summed across the wave, then added to per-workgroup float64 sums in LDS.
__global__ void __launch_bounds__(256)
linear_field_table_vjp_kernel(const float2 *__restrict__ wk, const float2 *__restrict__ gk, HalfGrid hg, Tabs t,
                              float3 box, float3 nm, float mcells, float vol, const float *__restrict__ ktab,
                              const float *__restrict__ pktab, int ntab, float scale, int nz,
                              double *__restrict__ partial) {
  extern __shared__ double tab_acc[];
  for (int e = threadIdx.x; e < ntab; e += blockDim.x) tab_acc[e] = 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const uint32_t stride = gridDim.x * blockDim.x;
  // every lane of a wave runs the same number of iterations (the bound is
  // rounded up to whole grid strides) so the cross-lane sums see full waves
  const uint32_t iters = (hg.H + stride - 1) / stride;
  for (uint32_t it = 0; it < iters; ++it) {
    const uint32_t e = it * stride + blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = e < hg.H;
    int seg = 1;
    double lo = 0.0, hi = 0.0;
    if (valid) {
      uint32_t i, j, k;
      hg.coords(e, i, j, k);
      float ax = t.k[i] / box.x * nm.x, ay = t.k[t.oy + j] / box.y * nm.y, az = t.k[t.oz + k] / box.z * nm.z;
      float kmesh = sqrtf((ax * ax + ay * ay) + az * az);
      float pk = interp(kmesh, ktab, pktab, ntab);
      if (pk > 0.0f) {
        float c = sqrtf(pk * mcells / vol) * scale;
        const float2 w = wk[e], g = gk[e];
        const double mult = (k > 0 && 2 * int(k) != nz) ? 2.0 : 1.0;
        const double val = mult * (double(w.x) * double(g.x) + double(w.y) * double(g.y)) * double(c) /
                           (2.0 * double(pk));
        int lo_i = 0, hi_i = ntab;
        while (lo_i < hi_i) {
          int mid = (lo_i + hi_i) >> 1;
          if (ktab[mid] <= kmesh) lo_i = mid + 1; else hi_i = mid;
        }
        seg = lo_i < 1 ? 1 : (lo_i > ntab - 1 ? ntab - 1 : lo_i);
        const float dx = ktab[seg] - ktab[seg - 1];
        if (kmesh < ktab[0]) {
          lo = val;                     // f = fp[0]
        } else if (kmesh > ktab[ntab - 1]) {
          hi = val;                     // f = fp[n-1]
        } else if (fabsf(dx) <= 0.0f) {
          lo = val;
        } else {
          const double tt = double((kmesh - ktab[seg - 1]) / dx);
          lo = val * (1.0 - tt);
          hi = val * tt;
        }
      }
    }
    uint64_t todo = __ballot(valid);
    while (todo) {
      const int leader = __ffsll((unsigned long long)todo) - 1;
      const int ls = __shfl(seg, leader);
      const bool mine = valid && seg == ls;
      todo &= ~uint64_t(__ballot(mine));
      double a = mine ? lo : 0.0, b = mine ? hi : 0.0;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
      }
      if (lane == leader) {
        atomicAdd(&tab_acc[ls - 1], a);
        atomicAdd(&tab_acc[ls], b);
      }
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ntab; e += blockDim.x) partial[size_t(blockIdx.x) * ntab + e] = tab_acc[e];
}

__global__ void table_reduce_kernel(const double *__restrict__ partial, int blocks, int ntab,
                                    double *__restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ntab) return;
  double s = 0.0;
  for (int b = 0; b < blocks; ++b) s += partial[size_t(b) * ntab + e];
  out[e] = s;
}

__global__ void __launch_bounds__(256)
shear6_kernel(const float2 *__restrict__ dk, float2 *__restrict__ out, HalfGrid hg, Tabs t, float scale) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < hg.H; e += gridDim.x * blockDim.x) {
    uint32_t i, j, k;
    hg.coords(e, i, j, k);
    const uint32_t jy = t.oy + j, kz = t.oz + k;
    float kk = (t.k2[i] + t.k2[jy]) + t.k2[kz];
    float G = green(kk) * scale;
    float2 v = dk[e];
    float pr = v.x * G, pi = v.y * G;
    // (i g_a)(i g_b) = -g_a g_b
    const float g[3] = {t.g[i], t.g[jy], t.g[kz]};
    const int A[6] = {0, 1, 2, 0, 0, 1};
    const int B[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      float c = -(g[A[m]] * g[B[m]]);
      out[size_t(m) * hg.H + e] = make_float2(c * pr, c * pi);
    }
  }
}

__global__ void __launch_bounds__(256)
lpt2_source_kernel(const float *__restrict__ phi, uint32_t m, float *__restrict__ out) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    float s00 = phi[e], s11 = phi[size_t(m) + e], s22 = phi[2 * size_t(m) + e];
    float s01 = phi[3 * size_t(m) + e], s02 = phi[4 * size_t(m) + e], s12 = phi[5 * size_t(m) + e];
    // pm.py:92-109 accumulation order
    float d = 0.0f + s00 * 0.0f;
    float acc = s00;
    d = d - s01 * s01;
    d = d - s02 * s02;
    d = d + s11 * acc;
    acc = acc + s11;
    d = d - s12 * s12;
    d = d + s22 * acc;
    out[e] = d;
  }
}

// Transpose of force_kernel as a real-linear map (the multiplier of the
// forward pass is -i a_d G; its adjoint multiplier is the conjugate, +i a_d G):
//   out = scale * sum_d (i a_d) G Lr [C] in_d.
__global__ void __launch_bounds__(256)
force_adjoint_kernel(const float2 *__restrict__ in3, float2 *__restrict__ out, HalfGrid hg, Tabs t, float r2,
                     float scale, int deconv) {
  // blockIdx.y: batch of equal shapes (in stride 3 H, out H)
  in3 += 3 * size_t(blockIdx.y) * hg.H;
  out += size_t(blockIdx.y) * hg.H;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < hg.H; e += gridDim.x * blockDim.x) {
    uint32_t i, j, k;
    hg.coords(e, i, j, k);
    const uint32_t jy = t.oy + j, kz = t.oz + k;
    float kk = (t.k2[i] + t.k2[jy]) + t.k2[kz];
    float f = green(kk);
    if (r2 != 0.0f) {
      float kke = (t.k2e[i] + t.k2e[jy]) + t.k2e[kz];
      f = f * expf(-kke * r2);
    }
    if (deconv) {
      float p = (t.s[i] * t.s[jy]) * t.s[kz];
      f = f / (p * p);
    }
    f = f * scale;
    const float2 a = in3[e], b = in3[size_t(hg.H) + e], c = in3[2 * size_t(hg.H) + e];
    const float gx = t.g[i] * f, gy = t.g[jy] * f, gz = t.g[kz] * f;
    // i g (re + i im) = -g im + i g re
    float re = -(gx * a.y) - gy * b.y - gz * c.y;
    float im = gx * a.x + gy * b.x + gz * c.x;
    out[e] = make_float2(re, im);
  }
}

// Adjoint of shear6_kernel: the six multipliers -g_a g_b G are real and even,
// so the map is its own transpose: out = scale * sum_m (-g_a g_b) G in_m.
__global__ void __launch_bounds__(256)
shear6_adjoint_kernel(const float2 *__restrict__ in6, float2 *__restrict__ out, HalfGrid hg, Tabs t, float scale) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < hg.H; e += gridDim.x * blockDim.x) {
    uint32_t i, j, k;
    hg.coords(e, i, j, k);
    const uint32_t jy = t.oy + j, kz = t.oz + k;
    float kk = (t.k2[i] + t.k2[jy]) + t.k2[kz];
    float G = green(kk) * scale;
    const float g[3] = {t.g[i], t.g[jy], t.g[kz]};
    const int A[6] = {0, 1, 2, 0, 0, 1};
    const int B[6] = {0, 1, 2, 1, 2, 2};
    float re = 0.0f, im = 0.0f;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const float c = -(g[A[m]] * g[B[m]]) * G;
      const float2 v = in6[size_t(m) * hg.H + e];
      re += c * v.x;
      im += c * v.y;
    }
    out[e] = make_float2(re, im);
  }
}

// Vector-Jacobian product of lpt2_source_kernel: given dL/d(delta2) per cell,
// write dL/d(phi_m) for the six shear meshes.
__global__ void __launch_bounds__(256)
lpt2_source_vjp_kernel(const float *__restrict__ phi, const float *__restrict__ gbar, uint32_t m,
                       float *__restrict__ out) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    const float s00 = phi[e], s11 = phi[size_t(m) + e], s22 = phi[2 * size_t(m) + e];
    const float s01 = phi[3 * size_t(m) + e], s02 = phi[4 * size_t(m) + e], s12 = phi[5 * size_t(m) + e];
    const float g = gbar[e];
    out[e] = g * (s11 + s22);
    out[size_t(m) + e] = g * (s00 + s22);
    out[2 * size_t(m) + e] = g * (s00 + s11);
    out[3 * size_t(m) + e] = -2.0f * g * s01;
    out[4 * size_t(m) + e] = -2.0f * g * s02;
    out[5 * size_t(m) + e] = -2.0f * g * s12;
  }
}

// Full c2c spectrum [nx][ny][nz] -> Hermitian part on the half grid:
// half(k) = (X(k) + conj(X(-k))) / 2.  The C2R of that half equals
// Re(ifftn(X)) * M for any X.
__global__ void __launch_bounds__(256)
full_to_half_kernel(const float2 *__restrict__ full, float2 *__restrict__ half, HalfGrid hg, uint32_t nx,
                    uint32_t nz) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < hg.H; e += gridDim.x * blockDim.x) {
    uint32_t i, j, k;
    hg.coords(e, i, j, k);
    const uint32_t mi = i ? nx - i : 0, mj = j ? hg.ny - j : 0, mk = k ? nz - k : 0;
    const float2 a = full[(size_t(i) * hg.ny + j) * nz + k];
    const float2 b = full[(size_t(mi) * hg.ny + mj) * nz + mk];
    half[e] = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
  }
}

// Half spectrum of a real field -> full c2c spectrum using X(-k) = conj(X(k)).
__global__ void __launch_bounds__(256)
half_to_full_kernel(const float2 *__restrict__ half, float2 *__restrict__ full, uint32_t M, uint32_t nx,
                    uint32_t ny, uint32_t nz, FastDiv div_nz, FastDiv div_ny) {
  const uint32_t nzh = nz / 2 + 1;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < M; e += gridDim.x * blockDim.x) {
    uint32_t r = div_nz.div(e);
    uint32_t k = e - r * nz;
    uint32_t i = div_ny.div(r);
    uint32_t j = r - i * ny;
    float2 v;
    if (k < nzh) {
      v = half[(size_t(i) * ny + j) * nzh + k];
    } else {
      const uint32_t mi = i ? nx - i : 0, mj = j ? ny - j : 0;
      const float2 b = half[(size_t(mi) * ny + mj) * nzh + (nz - k)];
      v = make_float2(b.x, -b.y);
    }
    full[e] = v;
  }
}

// ------------------------------------------------------------- host side

int kspace_force(hipStream_t s, int nx, int ny, int nz, const void *dk, void *out3, float r_split, float scale,
                 uint32_t flags) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(dk && out3, "null spectrum pointer");
  JPM_REQUIRE(3 * half_cells(nx, ny, nz) <= (int64_t(1) << 32), "spectrum too large");
  Tabs t;
  rc = get_tabs(s, nx, ny, nz, flags, t);
  if (rc) return rc;
  HalfGrid hg = make_half(nx, ny, nz);
  {
    JPM_TIMED(s, "force_kernel");
    hipLaunchKernelGGL(force_kernel, dim3(grid_for(hg.H, 256, 256 * 128)), dim3(256), 0, s,
                       static_cast<const float2 *>(dk), static_cast<float2 *>(out3), hg, t, r_split * r_split,
                       scale, int((flags & JPM_K_DECONV) != 0));
  }
  JPM_LAUNCHED("force_kernel");
  return JPM_OK;
}

int kspace_potential_batch(hipStream_t s, int nx, int ny, int nz, int batch, const void *dk, void *out,
                           float r_split, float scale, const float *scale_dev, uint32_t flags) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(dk && out, "null spectrum pointer");
  JPM_REQUIRE(batch >= 1 && batch <= 65535, "batch of %d spectra", batch);
  Tabs t;
  rc = get_tabs(s, nx, ny, nz, flags, t);
  if (rc) return rc;
  HalfGrid hg = make_half(nx, ny, nz);
  {
    JPM_TIMED(s, "potential_kernel");
    hipLaunchKernelGGL(potential_kernel, dim3(grid_for(hg.H, 256, 256 * 128 / batch + 1), batch), dim3(256), 0, s,
                       static_cast<const float2 *>(dk), static_cast<float2 *>(out), hg, t, r_split * r_split, scale,
                       scale_dev, int((flags & JPM_K_DECONV) != 0));
  }
  JPM_LAUNCHED("potential_kernel");
  return JPM_OK;
}

int fft_r2c(hipStream_t s, int nx, int ny, int nz, int batch, const float *in, void *out);
int fft_c2r(hipStream_t s, int nx, int ny, int nz, int batch, void *in, float *out);
bool native_fft_ok(int nx, int ny, int nz);
int native_potential(hipStream_t s, int nx, int ny, int nz, int batch, const float *in, void *spec, float *out,
                     const Tabs &t, float r2, float scale, const float *scale_dev, int deconv,
                     const float *g3 = nullptr, size_t chan = 0, const FaceCells *faces = nullptr);
int fd4_divergence_batch(hipStream_t s, int nx, int ny, int nz, int batch, const float *g3, size_t chan_stride,
                         float scale, float *out);

// out = C2R(potential(R2C(in))) for a batch of equal meshes (in and out may
// be the same array): the forward transform, potential_kernel's multiplier
// and the inverse transform.  spec and spec2 hold H complex values per mesh;
// power-of-two meshes apply the multiplier inside the FFT's x pass and use
// spec only, other shapes run hipFFT, potential_kernel into spec2, hipFFT.
// The shapes on which the potential solves run the library's own FFT, and
// on which they can read a paint's face cells kept apart (FaceCells): one
// mesh of full 16^3 tiles.
bool potential_faces_ok(int nx, int ny, int nz) {
  return native_fft_ok(nx, ny, nz) && nx >= 2 && nx % kTile == 0 && ny % kTile == 0 && nz % kTile == 0;
}

int potential_solve(hipStream_t s, int nx, int ny, int nz, int batch, const float *in, void *spec, void *spec2,
                    float *out, float r_split, float scale, const float *scale_dev, uint32_t flags,
                    const FaceCells *faces) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(in && spec && spec2 && out, "null pointer argument");
  JPM_REQUIRE(batch >= 1 && batch <= 65535, "batch of %d meshes", batch);
  const bool aligned = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 7u) == 0;
  JPM_REQUIRE(!faces || (potential_faces_ok(nx, ny, nz) && aligned && batch == 1),
              "face cells apart need the library's FFT on one mesh of full tiles");
  if (native_fft_ok(nx, ny, nz) && nx >= 2 && aligned) {
    Tabs t;
    rc = get_tabs(s, nx, ny, nz, flags, t);
    if (rc) return rc;
    JPM_TIMED(s, "potential_solve (5 passes)");
    return native_potential(s, nx, ny, nz, batch, in, spec, out, t, r_split * r_split, scale, scale_dev,
                            int((flags & JPM_K_DECONV) != 0), nullptr, 0, faces);
  }
  if ((rc = fft_r2c(s, nx, ny, nz, batch, in, spec))) return rc;
  if ((rc = kspace_potential_batch(s, nx, ny, nz, batch, spec, spec2, r_split, scale, scale_dev, flags))) return rc;
  return fft_c2r(s, nx, ny, nz, batch, spec2, out);
}

// out = C2R(potential(R2C(sum_d D_d g_d))) for a batch of equal meshes: the
// force VJP's density adjoint from the painted cotangent g (channel d of
// item b at g3 + d chan + b nx ny nz).  Power-of-two meshes form the
// divergence inside the R2C's first pass; other shapes run
// fd4_divergence_kernel into rho (which may be out), then potential_solve.
int potential_solve_div(hipStream_t s, int nx, int ny, int nz, int batch, const float *g3, size_t chan, float *rho,
                        void *spec, void *spec2, float *out, float r_split, float scale, const float *scale_dev,
                        uint32_t flags) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(g3 && rho && spec && spec2 && out, "null pointer argument");
  JPM_REQUIRE(batch >= 1 && batch <= 65535, "batch of %d meshes", batch);
  const bool aligned =
      ((reinterpret_cast<uintptr_t>(g3) | reinterpret_cast<uintptr_t>(out) | uintptr_t(chan * sizeof(float))) & 7u) == 0;
  if (native_fft_ok(nx, ny, nz) && nx >= 2 && aligned) {
    Tabs t;
    rc = get_tabs(s, nx, ny, nz, flags, t);
    if (rc) return rc;
    JPM_TIMED(s, "potential_solve[div] (5 passes)");
    return native_potential(s, nx, ny, nz, batch, nullptr, spec, out, t, r_split * r_split, scale, scale_dev,
                            int((flags & JPM_K_DECONV) != 0), g3, chan);
  }
  if ((rc = fd4_divergence_batch(s, nx, ny, nz, batch, g3, chan, 1.0f, rho))) return rc;
  return potential_solve(s, nx, ny, nz, batch, rho, spec, spec2, out, r_split, scale, scale_dev, flags);
}

int kspace_potential(hipStream_t s, int nx, int ny, int nz, const void *dk, void *out, float r_split, float scale,
                     uint32_t flags) {
  return kspace_potential_batch(s, nx, ny, nz, 1, dk, out, r_split, scale, nullptr, flags);
}

int kspace_compensate(hipStream_t s, int nx, int ny, int nz, void *fk, float scale) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(fk, "null spectrum pointer");
  Tabs t;
  rc = get_tabs(s, nx, ny, nz, 0, t);
  if (rc) return rc;
  HalfGrid hg = make_half(nx, ny, nz);
  hipLaunchKernelGGL(compensate_kernel, dim3(grid_for(hg.H, 256, 256 * 128)), dim3(256), 0, s,
                     static_cast<float2 *>(fk), hg, t, scale);
  JPM_LAUNCHED("compensate_kernel");
  return JPM_OK;
}

int kspace_linear_field(hipStream_t s, int nx, int ny, int nz, void *fk, float Lx, float Ly, float Lz,
                        const float *ktab, const float *pktab, int ntab, float scale) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(fk && ktab && pktab, "null pointer argument");
  JPM_REQUIRE(ntab >= 2, "power-spectrum table needs at least 2 entries (got %d)", ntab);
  JPM_REQUIRE(Lx > 0 && Ly > 0 && Lz > 0, "box size must be positive");
  Tabs t;
  rc = get_tabs(s, nx, ny, nz, 0, t);
  if (rc) return rc;
  HalfGrid hg = make_half(nx, ny, nz);
  const float3 box = make_float3(Lx, Ly, Lz);
  const float3 nm = make_float3(float(nx), float(ny), float(nz));
  const float mcells = float(cells(nx, ny, nz));
  const float vol = float(double(Lx) * double(Ly) * double(Lz));
  {
    JPM_TIMED(s, "linear_field_kernel");
    hipLaunchKernelGGL(linear_field_kernel, dim3(grid_for(hg.H, 256, 256 * 128)), dim3(256), 0, s,
                       static_cast<float2 *>(fk), hg, t, box, nm, mcells, vol, ktab, pktab, ntab, scale,
                       static_cast<const float *>(nullptr));
  }
  JPM_LAUNCHED("linear_field_kernel");
  return JPM_OK;
}

// Batched forms over `batch` spectra of one shape (stride H; the force
// passes' three-component spectra stride 3 H).
int kspace_force_batch(hipStream_t s, int nx, int ny, int nz, int batch, const void *dk, void *out3, float scale,
                       uint32_t flags) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(dk && out3 && batch >= 1 && batch <= 65535, "bad batched force pass arguments");
  Tabs t;
  rc = get_tabs(s, nx, ny, nz, flags, t);
  if (rc) return rc;
  HalfGrid hg = make_half(nx, ny, nz);
  JPM_TIMED(s, "force_kernel");
  hipLaunchKernelGGL(force_kernel, dim3(grid_for(hg.H, 256, 256 * 128 / batch + 1), batch), dim3(256), 0, s,
                     static_cast<const float2 *>(dk), static_cast<float2 *>(out3), hg, t, 0.0f, scale,
                     int((flags & JPM_K_DECONV) != 0));
  JPM_LAUNCHED("force_kernel");
  return JPM_OK;
}

int kspace_force_adjoint_batch(hipStream_t s, int nx, int ny, int nz, int batch, const void *in3, void *out,
                               float scale, uint32_t flags) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(in3 && out && batch >= 1 && batch <= 65535, "bad batched force adjoint arguments");
  Tabs t;
  rc = get_tabs(s, nx, ny, nz, flags, t);
  if (rc) return rc;
  HalfGrid hg = make_half(nx, ny, nz);
  JPM_TIMED(s, "force_adjoint_kernel");
  hipLaunchKernelGGL(force_adjoint_kernel, dim3(grid_for(hg.H, 256, 256 * 128 / batch + 1), batch), dim3(256), 0, s,
                     static_cast<const float2 *>(in3), static_cast<float2 *>(out), hg, t, 0.0f, scale,
                     int((flags & JPM_K_DECONV) != 0));
  JPM_LAUNCHED("force_adjoint_kernel");
  return JPM_OK;
}

int kspace_linear_field_batch(hipStream_t s, int nx, int ny, int nz, int batch, void *fk, const float *boxes,
                              const float *ktab, const float *pktabs, int ntab, float scale) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(fk && boxes && ktab && pktabs && batch >= 1 && batch <= 65535, "bad batched linear field arguments");
  JPM_REQUIRE(ntab >= 2, "power-spectrum table needs at least 2 entries (got %d)", ntab);
  Tabs t;
  rc = get_tabs(s, nx, ny, nz, 0, t);
  if (rc) return rc;
  HalfGrid hg = make_half(nx, ny, nz);
  const float3 nm = make_float3(float(nx), float(ny), float(nz));
  JPM_TIMED(s, "linear_field_kernel");
  hipLaunchKernelGGL(linear_field_kernel, dim3(grid_for(hg.H, 256, 256 * 128 / batch + 1), batch), dim3(256), 0, s,
                     static_cast<float2 *>(fk), hg, t, make_float3(1.0f, 1.0f, 1.0f), nm,
                     float(cells(nx, ny, nz)), 1.0f, ktab, pktabs, ntab, scale, boxes);
  JPM_LAUNCHED("linear_field_kernel");
  return JPM_OK;
}

constexpr int kTableVjpBlocks = 256;
constexpr int kTableVjpMax = 8192;   // float64 LDS sums per workgroup (64 KiB)

size_t linear_field_table_vjp_ws(int ntab) {
  return ntab < 2 ? 0 : size_t(kTableVjpBlocks) * size_t(ntab) * sizeof(double);
}

int kspace_linear_field_table_vjp(hipStream_t s, int nx, int ny, int nz, const void *wk, const void *gk, float Lx,
                                  float Ly, float Lz, const float *ktab, const float *pktab, int ntab, float scale,
                                  double *out, void *ws, size_t ws_bytes) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(wk && gk && ktab && pktab && out, "null pointer argument");
  JPM_REQUIRE(ntab >= 2 && ntab <= kTableVjpMax, "power-spectrum table of %d entries (2..%d)", ntab,
              kTableVjpMax);
  JPM_REQUIRE(Lx > 0 && Ly > 0 && Lz > 0, "box size must be positive");
  JPM_REQUIRE(ws && ws_bytes >= linear_field_table_vjp_ws(ntab), "workspace of %zu bytes, need %zu", ws_bytes,
              linear_field_table_vjp_ws(ntab));
  Tabs t;
  rc = get_tabs(s, nx, ny, nz, 0, t);
  if (rc) return rc;
  HalfGrid hg = make_half(nx, ny, nz);
  const float3 box = make_float3(Lx, Ly, Lz);
  const float3 nm = make_float3(float(nx), float(ny), float(nz));
  const float mcells = float(cells(nx, ny, nz));
  const float vol = float(double(Lx) * double(Ly) * double(Lz));
  double *partial = static_cast<double *>(ws);
  {
    JPM_TIMED(s, "linear_field_table_vjp_kernel");
    hipLaunchKernelGGL(linear_field_table_vjp_kernel, dim3(kTableVjpBlocks), dim3(256), size_t(ntab) * sizeof(double),
                       s, static_cast<const float2 *>(wk), static_cast<const float2 *>(gk), hg, t, box, nm, mcells,
                       vol, ktab, pktab, ntab, scale, nz, partial);
    JPM_LAUNCHED("linear_field_table_vjp_kernel");
  }
  hipLaunchKernelGGL(table_reduce_kernel, dim3((ntab + 255) / 256), dim3(256), 0, s, partial, kTableVjpBlocks, ntab,
                     out);
  JPM_LAUNCHED("table_reduce_kernel");
  return JPM_OK;
}

int kspace_shear6(hipStream_t s, int nx, int ny, int nz, const void *dk, void *out6, float scale,
                  uint32_t flags) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(dk && out6, "null spectrum pointer");
  JPM_REQUIRE(6 * half_cells(nx, ny, nz) <= (int64_t(1) << 32), "spectrum too large");
  Tabs t;
  rc = get_tabs(s, nx, ny, nz, flags, t);
  if (rc) return rc;
  HalfGrid hg = make_half(nx, ny, nz);
  {
    JPM_TIMED(s, "shear6_kernel");
    hipLaunchKernelGGL(shear6_kernel, dim3(grid_for(hg.H, 256, 256 * 128)), dim3(256), 0, s,
                       static_cast<const float2 *>(dk), static_cast<float2 *>(out6), hg, t, scale);
  }
  JPM_LAUNCHED("shear6_kernel");
  return JPM_OK;
}

int lpt2_source(hipStream_t s, int64_t m, const float *phi6, float *delta2) {
  JPM_REQUIRE(m > 0 && m < (int64_t(1) << 32), "mesh size %lld out of range", (long long)m);  // size_t channel offsets
  JPM_REQUIRE(phi6 && delta2, "null pointer argument");
  hipLaunchKernelGGL(lpt2_source_kernel, dim3(grid_for(m, 256, 256 * 128)), dim3(256), 0, s, phi6,
                     uint32_t(m), delta2);
  JPM_LAUNCHED("lpt2_source_kernel");
  return JPM_OK;
}

int kspace_force_adjoint(hipStream_t s, int nx, int ny, int nz, const void *in3, void *out, float r_split,
                         float scale, uint32_t flags) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(in3 && out, "null spectrum pointer");
  JPM_REQUIRE(3 * half_cells(nx, ny, nz) <= (int64_t(1) << 32), "spectrum too large");
  Tabs t;
  rc = get_tabs(s, nx, ny, nz, flags, t);
  if (rc) return rc;
  HalfGrid hg = make_half(nx, ny, nz);
  {
    JPM_TIMED(s, "force_adjoint_kernel");
    hipLaunchKernelGGL(force_adjoint_kernel, dim3(grid_for(hg.H, 256, 256 * 128)), dim3(256), 0, s,
                       static_cast<const float2 *>(in3), static_cast<float2 *>(out), hg, t, r_split * r_split, scale,
                       int((flags & JPM_K_DECONV) != 0));
  }
  JPM_LAUNCHED("force_adjoint_kernel");
  return JPM_OK;
}

int kspace_shear6_adjoint(hipStream_t s, int nx, int ny, int nz, const void *in6, void *out, float scale,
                          uint32_t flags) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(in6 && out, "null spectrum pointer");
  JPM_REQUIRE(6 * half_cells(nx, ny, nz) <= (int64_t(1) << 32), "spectrum too large");
  Tabs t;
  rc = get_tabs(s, nx, ny, nz, flags, t);
  if (rc) return rc;
  HalfGrid hg = make_half(nx, ny, nz);
  hipLaunchKernelGGL(shear6_adjoint_kernel, dim3(grid_for(hg.H, 256, 256 * 128)), dim3(256), 0, s,
                     static_cast<const float2 *>(in6), static_cast<float2 *>(out), hg, t, scale);
  JPM_LAUNCHED("shear6_adjoint_kernel");
  return JPM_OK;
}

int lpt2_source_vjp(hipStream_t s, int64_t m, const float *phi6, const float *gbar, float *out6) {
  JPM_REQUIRE(m > 0 && m < (int64_t(1) << 32), "mesh size %lld out of range", (long long)m);  // size_t channel offsets
  JPM_REQUIRE(phi6 && gbar && out6, "null pointer argument");
  hipLaunchKernelGGL(lpt2_source_vjp_kernel, dim3(grid_for(m, 256, 256 * 128)), dim3(256), 0, s, phi6, gbar,
                     uint32_t(m), out6);
  JPM_LAUNCHED("lpt2_source_vjp_kernel");
  return JPM_OK;
}

// Complex arrays [batch][M] -> real fields [2 batch][M]: x = Re z, y =
// sign Im z (sign -1 conjugates, for the inverse transform).
__global__ void __launch_bounds__(256)
c2c_split_kernel(const float2 *__restrict__ z, float *__restrict__ xy, uint32_t M, float sign) {
  const float2 *zb = z + size_t(blockIdx.y) * M;
  float *x = xy + 2 * size_t(blockIdx.y) * M;
  float *y = x + M;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < M; e += gridDim.x * blockDim.x) {
    const float2 v = zb[e];
    x[e] = v.x;
    y[e] = sign * v.y;
  }
}

// Half spectra X, Y of the pair (x, y) of array b -> full spectrum of
// x + iy: Z(k) = X(k) + i Y(k), with X(k) = conj X(-k) above nz/2; the
// imaginary part times sign (-1: the conjugate, for the inverse).
__global__ void __launch_bounds__(256)
c2c_combine_kernel(const float2 *__restrict__ xyh, float2 *__restrict__ out, uint32_t M, uint32_t H, uint32_t nx,
                   uint32_t ny, uint32_t nz, FastDiv div_nz, FastDiv div_ny, float sign) {
  const float2 *X = xyh + 2 * size_t(blockIdx.y) * H;
  const float2 *Y = X + H;
  float2 *o = out + size_t(blockIdx.y) * M;
  const uint32_t nzh = nz / 2 + 1;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < M; e += gridDim.x * blockDim.x) {
    const uint32_t r = div_nz.div(e);
    const uint32_t k = e - r * nz;
    const uint32_t i = div_ny.div(r);
    const uint32_t j = r - i * ny;
    float2 a, b;
    if (k < nzh) {
      const size_t h = (size_t(i) * ny + j) * nzh + k;
      a = X[h];
      b = Y[h];
    } else {
      const uint32_t mi = i ? nx - i : 0, mj = j ? ny - j : 0;
      const size_t h = (size_t(mi) * ny + mj) * nzh + (nz - k);
      a = make_float2(X[h].x, -X[h].y);
      b = make_float2(Y[h].x, -Y[h].y);
    }
    o[e] = make_float2(a.x - b.y, sign * (a.y + b.x));
  }
}

int c2c_split(hipStream_t s, int64_t m, int batch, const void *z, float *xy, float sign) {
  JPM_REQUIRE(m > 0 && m < (int64_t(1) << 32) && batch >= 1 && batch <= 65535, "c2c split of %d x %lld", batch,
              (long long)m);
  hipLaunchKernelGGL(c2c_split_kernel, dim3(grid_for(m, 256, 256 * 64 / batch + 1), batch), dim3(256), 0, s,
                     static_cast<const float2 *>(z), xy, uint32_t(m), sign);
  JPM_LAUNCHED("c2c_split_kernel");
  return JPM_OK;
}

int c2c_combine(hipStream_t s, int nx, int ny, int nz, int batch, const void *xyh, void *out, float sign) {
  const int64_t m = cells(nx, ny, nz);
  JPM_REQUIRE(m < (int64_t(1) << 32) && batch >= 1 && batch <= 65535, "c2c combine of %d x %lld", batch,
              (long long)m);
  hipLaunchKernelGGL(c2c_combine_kernel, dim3(grid_for(m, 256, 256 * 64 / batch + 1), batch), dim3(256), 0, s,
                     static_cast<const float2 *>(xyh), static_cast<float2 *>(out), uint32_t(m),
                     uint32_t(half_cells(nx, ny, nz)), uint32_t(nx), uint32_t(ny), uint32_t(nz),
                     FastDiv(uint32_t(nz)), FastDiv(uint32_t(ny)), sign);
  JPM_LAUNCHED("c2c_combine_kernel");
  return JPM_OK;
}

int spectrum_full_to_half(hipStream_t s, int nx, int ny, int nz, const void *full, void *half) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(full && half, "null spectrum pointer");
  HalfGrid hg = make_half(nx, ny, nz);
  hipLaunchKernelGGL(full_to_half_kernel, dim3(grid_for(hg.H, 256, 256 * 128)), dim3(256), 0, s,
                     static_cast<const float2 *>(full), static_cast<float2 *>(half), hg, uint32_t(nx), uint32_t(nz));
  JPM_LAUNCHED("full_to_half_kernel");
  return JPM_OK;
}

int spectrum_half_to_full(hipStream_t s, int nx, int ny, int nz, const void *half, void *full) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(full && half, "null spectrum pointer");
  const int64_t m = cells(nx, ny, nz);
  hipLaunchKernelGGL(half_to_full_kernel, dim3(grid_for(m, 256, 256 * 128)), dim3(256), 0, s,
                     static_cast<const float2 *>(half), static_cast<float2 *>(full), uint32_t(m), uint32_t(nx),
                     uint32_t(ny), uint32_t(nz), FastDiv(uint32_t(nz)), FastDiv(uint32_t(ny)));
  JPM_LAUNCHED("half_to_full_kernel");
  return JPM_OK;
}

}  // namespace jpm
