// Per-axis Fourier multiplier tables and the potential multiplier, shared by
// the k-space passes (kspace.hip) and the FFT pass that applies the
// potential between its forward and inverse column transforms (fftk.hip).
#pragma once

#include "common.hpp"

namespace jpm {

// Per-axis tables, each of length nx + ny + nzh (x entries, then y, then z):
//   k    : wave number 2*pi*fftfreq(n) in radians per cell (fftk)
//   k2   : k^2, or (k sinc(k/2pi))^2 with JPM_K_LAPL_FD (invlaplace_kernel)
//   k2e  : k^2 (longrange_kernel always uses the plain k^2)
//   g    : gradient_kernel coefficient a, the kernel being i*a
//   s    : sinc(k/2pi) (cic_compensation = (sx sy sz)^-2)
struct Tabs {
  const float *k, *k2, *k2e, *g, *s;
  uint32_t oy, oz;  // offsets of the y and z entries
};

// The cached tables of a mesh shape and flags (JPM_K_GRAD_EXACT,
// JPM_K_LAPL_FD select the variants).
int get_tabs(hipStream_t st, int nx, int ny, int nz, uint32_t flags, Tabs &t);

// invlaplace_kernel (kernels.py:90-92): -1/kk, 0 at kk == 0.
__device__ __forceinline__ float green(float kk) { return kk == 0.0f ? 0.0f : -(1.0f / kk); }

// The potential multiplier at mode (i, jy, kz) (table indices, jy and kz
// already offset): scale * G(k) [* exp(-k^2 r2)] [/ (sx sy sz)^2]
// (pm.py:49-52 with kernels.py:69-115).
__device__ __forceinline__ float potential_factor(const Tabs &t, uint32_t i, uint32_t jy, uint32_t kz, float r2,
                                                  int deconv, float scale) {
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
  return f * scale;
}

}  // namespace jpm
