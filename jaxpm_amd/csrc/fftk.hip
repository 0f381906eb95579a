// Native 3-D real-to-complex / complex-to-real FFTs for power-of-two meshes,
// the transforms of jaxdecomp.pfft3d / pifft3d (distributed.py:37-42) with
// hipFFT's conventions: unnormalised, e^{-i} forward, half spectrum
// [nx][ny][nz/2+1].
//
// Each pass reads and writes the data once:
//   R2C  z: every row of nz reals as nz/2 complex points (even, odd
//           samples), an nz/2-point FFT, then the split into the nz/2 + 1
//           half-spectrum values;  y, x: in-place column FFTs over kz blocks.
//   C2R  x, y: in-place inverse column FFTs;  z: the merge of each
//           half-spectrum row into nz/2 complex points, an inverse FFT, the
//           real row (imaginary parts of kz = 0 and nz/2 dropped, as irfft).
// A workgroup stages up to 4096 complex points (KB columns of length L) in
// LDS and runs a Stockham FFT of radix-16 stages (then one radix-8, 4 or 2
// stage) with twiddles from a table computed in double precision.
#include <mutex>
#include <vector>

#include "kspace.hpp"
#include "tiles.hpp"

namespace jpm {

constexpr int kFftThreads = 512;
constexpr int kFftPoints = 4096;   // complex points per workgroup
constexpr int kFftMaxL = 1024;

// The transform lengths: powers of two 2 .. 1024 and 3 * 2^k, 6 .. 768.
constexpr bool fft_len(int L) {
  return L >= 2 && (((L & (L - 1)) == 0 && L <= kFftMaxL) ||
                    (L % 3 == 0 && ((L / 3) & (L / 3 - 1)) == 0 && L <= 3 * kFftMaxL / 4));
}

// exp(-2 pi i e / L) for e in [0, L), one table per length L = 2 .. 2048
// (powers of two, at offset L - 2 of one device buffer) and L = 6 .. 1536
// (3 * 2^k, at offset kTwPow2 + L - 6 after them).
constexpr int kTwPow2 = 4 * kFftMaxL - 2;
constexpr int tw_off(int L) { return (L & (L - 1)) == 0 ? L - 2 : kTwPow2 + (L - 6); }

static std::mutex g_tw_mu;
static float2 *g_tw = nullptr;

static int twiddles(const float2 **out) {
  std::lock_guard<std::mutex> lk(g_tw_mu);
  if (!g_tw) {
    std::vector<float2> h;
    auto table = [&](int L) {
      for (int e = 0; e < L; ++e) {
        const double a = -2.0 * M_PI * double(e) / double(L);
        h.push_back(make_float2(float(std::cos(a)), float(std::sin(a))));
      }
    };
    for (int L = 2; L <= 2 * kFftMaxL; L *= 2) table(L);
    for (int L = 6; L <= 3 * kFftMaxL / 2; L *= 2) table(L);
    float *buf = nullptr;
    JPM_HIP(hipMalloc(&buf, h.size() * sizeof(float2)));
    const hipError_t e = hipMemcpy(buf, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(buf);
      return fail(JPM_E_HIP, "uploading FFT twiddles failed: %s", hipGetErrorString(e));
    }
    g_tw = reinterpret_cast<float2 *>(buf);
  }
  *out = g_tw;
  return JPM_OK;
}

void free_twiddles() {
  std::lock_guard<std::mutex> lk(g_tw_mu);
  if (g_tw) (void)hipFree(g_tw);
  g_tw = nullptr;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// 4-point DFT in place, e^{-i} (forward) or e^{+i} (INV).
template <bool INV>
__device__ __forceinline__ void dft4(float2 &a0, float2 &a1, float2 &a2, float2 &a3) {
  const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), d = csub(a1, a3);
  const float2 t3 = INV ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);   // d * (+-i)
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

// a * W_16^e (e^{-2 pi i e / 16}, conjugated for INV) for the constant
// exponents of the radix-8 and radix-16 butterflies.
template <bool INV, int E>
__device__ __forceinline__ float2 w16(float2 a) {
  constexpr float c1 = 0.92387953251128674f, s1 = 0.38268343236508977f, h = 0.70710678118654752f;
  constexpr int e = E & 15;
  float cr, ci;   // W = cr - i ci (forward)
  if constexpr (e == 0) { return a; }
  else if constexpr (e == 4) { return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x); }
  else if constexpr (e == 8) { return make_float2(-a.x, -a.y); }
  else if constexpr (e == 12) { return INV ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x); }
  else {
    constexpr float tc[16] = {1.0f, c1, h, s1, 0.0f, -s1, -h, -c1, -1.0f, -c1, -h, -s1, 0.0f, s1, h, c1};
    constexpr float ts[16] = {0.0f, s1, h, c1, 1.0f, c1, h, s1, 0.0f, -s1, -h, -c1, -1.0f, -c1, -h, -s1};
    cr = tc[e];
    ci = INV ? -ts[e] : ts[e];
    return make_float2(a.x * cr + a.y * ci, a.y * cr - a.x * ci);
  }
}

// R-point DFT of a[0..R) in place, natural order, R = 2, 3, 4, 8 or 16:
// a length-R transform as (R/4 or 2) x 4 with the W_R twiddles between.
template <int R, bool INV>
__device__ __forceinline__ void dftR(float2 *a) {
  if constexpr (R == 3) {
    // X1, X2 = a0 - (a1 + a2)/2 -+ i sin(2 pi / 3) (a1 - a2) (forward)
    constexpr float s3 = 0.86602540378443865f;
    const float2 t = cadd(a[1], a[2]), d = csub(a[1], a[2]);
    const float2 m = make_float2(a[0].x - 0.5f * t.x, a[0].y - 0.5f * t.y);
    const float2 r = INV ? make_float2(-s3 * d.y, s3 * d.x) : make_float2(s3 * d.y, -s3 * d.x);   // d * (+-i) s3
    a[0] = cadd(a[0], t);
    a[1] = cadd(m, r);
    a[2] = csub(m, r);
  } else if constexpr (R == 2) {
    const float2 b0 = cadd(a[0], a[1]), b1 = csub(a[0], a[1]);
    a[0] = b0;
    a[1] = b1;
  } else if constexpr (R == 4) {
    dft4<INV>(a[0], a[1], a[2], a[3]);
  } else if constexpr (R == 8) {
    // m = 2 m1 + m2 (m1 < 4, m2 < 2); q = q1 + 4 q2
    dft4<INV>(a[0], a[2], a[4], a[6]);
    dft4<INV>(a[1], a[3], a[5], a[7]);
    a[3] = w16<INV, 2>(a[3]);   // W_8^{m2 q1} = W_16^{2 m2 q1}
    a[5] = w16<INV, 4>(a[5]);
    a[7] = w16<INV, 6>(a[7]);
    float2 b[8];
#pragma unroll
    for (int q1 = 0; q1 < 4; ++q1) {
      b[q1] = cadd(a[2 * q1], a[2 * q1 + 1]);
      b[q1 + 4] = csub(a[2 * q1], a[2 * q1 + 1]);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] = b[q];
  } else {
    // m = 4 m1 + m2; q = q1 + 4 q2
#pragma unroll
    for (int m2 = 0; m2 < 4; ++m2) dft4<INV>(a[m2], a[4 + m2], a[8 + m2], a[12 + m2]);
    // y[m2][q1] = a[4 q1 + m2] now; twiddle by W_16^{m2 q1}
    a[5] = w16<INV, 1>(a[5]);
    a[9] = w16<INV, 2>(a[9]);
    a[13] = w16<INV, 3>(a[13]);
    a[6] = w16<INV, 2>(a[6]);
    a[10] = w16<INV, 4>(a[10]);
    a[14] = w16<INV, 6>(a[14]);
    a[7] = w16<INV, 3>(a[7]);
    a[11] = w16<INV, 6>(a[11]);
    a[15] = w16<INV, 9>(a[15]);
#pragma unroll
    for (int q1 = 0; q1 < 4; ++q1) dft4<INV>(a[4 * q1], a[4 * q1 + 1], a[4 * q1 + 2], a[4 * q1 + 3]);
    // X[q1 + 4 q2] = a[4 q1 + q2]: transpose the 4 x 4 block
    float2 b[16];
#pragma unroll
    for (int q1 = 0; q1 < 4; ++q1)
#pragma unroll
      for (int q2 = 0; q2 < 4; ++q2) b[q1 + 4 * q2] = a[4 * q1 + q2];
#pragma unroll
    for (int q = 0; q < 16; ++q) a[q] = b[q];
  }
}

// Radix of the next stage: 3 while the points left hold a factor 3, then
// 16 while at least 16 remain, then what is left (8, 4 or 2).
constexpr int stage_radix(int rem) { return rem % 3 == 0 ? 3 : rem >= 16 ? 16 : rem; }

// FFTs of KB columns of length L held in LDS as buf[l * PITCH + c]
// (c < KB), in place; L, KB and PITCH are compile-time, so every index is
// shifts and masks and the stage loop unrolls.  Stockham stage with span Ns
// and radix R: butterfly j of a column reads elements j + r L/R, twiddles
// them by W_L^{r k L/(Ns R)} (k = j mod Ns), and writes element
// (j / Ns) Ns R + k + r Ns.  Every thread keeps its butterflies in registers
// between the read and the write half of a stage.  L * KB <= 4096, so a
// thread holds at most 16 points.  tw is the length-L table in LDS.
template <int L, int KB, int PITCH, bool INV, int NS, int NT = kFftThreads>
__device__ __forceinline__ void fft_stage(float2 *buf, const float2 *tw) {
  constexpr int R = stage_radix(L / NS);
  constexpr int nb = L / R, step = L / (NS * R), total = nb * KB;
  constexpr int per = (total + NT - 1) / NT;
  float2 x[per][R];
  int dst[per], cc[per];
#pragma unroll
  for (int it = 0; it < per; ++it) {
    const int f = threadIdx.x + it * NT;
    if (total % NT == 0 || f < total) {
      const int c = f & (KB - 1), j = f / KB, k = j % NS;
      float2 a[R];
#pragma unroll
      for (int r = 0; r < R; ++r) a[r] = buf[(j + r * nb) * PITCH + c];
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[r * k * step];   // r k step < R NS step = L
          if (INV) w = conjf2(w);
          a[r] = cmul(a[r], w);
        }
      }
      dftR<R, INV>(a);
#pragma unroll
      for (int r = 0; r < R; ++r) x[it][r] = a[r];
      dst[it] = (j / NS) * NS * R + k;
      cc[it] = c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < per; ++it) {
    const int f = threadIdx.x + it * NT;
    if (total % NT == 0 || f < total) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(dst[it] + r * NS) * PITCH + cc[it]] = x[it][r];
    }
  }
  __syncthreads();
  constexpr int next = NS * R;
  if constexpr (next < L) fft_stage<L, KB, PITCH, INV, next, NT>(buf, tw);
}

template <int L, int KB, int PITCH, bool INV, int NT = kFftThreads>
__device__ __forceinline__ void lds_fft(float2 *buf, const float2 *tw) {
  if constexpr (L > 1) fft_stage<L, KB, PITCH, INV, 1, NT>(buf, tw);
}

// A twiddle table into LDS in two halves, so that its loads share their
// latency with the data loads: load() reads entries [off, off + N) of the
// device table into registers, store() writes them to LDS (the barrier after
// the data stores orders them before the first stage that reads them).
template <int N, int NT = kFftThreads>
struct TwStage {
  static constexpr int I = (N + NT - 1) / NT;
  float2 v[I];
  __device__ __forceinline__ void load(const float2 *__restrict__ tw_all, int off) {
#pragma unroll
    for (int i = 0; i < I; ++i) {
      const int e = int(threadIdx.x) + i * NT;
      v[i] = tw_all[off + (e < N ? e : N - 1)];
    }
  }
  __device__ __forceinline__ void store(float2 *tws) const {
#pragma unroll
    for (int i = 0; i < I; ++i) {
      const int e = int(threadIdx.x) + i * NT;
      if (e < N) tws[e] = v[i];
    }
  }
};

// Columns per workgroup: the largest power of two <= min(64, 4096 / L).
constexpr int pow2_floor(int v) { return v < 2 ? 1 : 2 * pow2_floor(v / 2); }
template <int L>
struct FftCols {
  static constexpr int KB = pow2_floor((kFftPoints / L) > 64 ? 64 : (kFftPoints / L));
};

// The KB columns of length L at kz0 .. kz0 + KB - 1 into buf[l * KB + c]:
// every load of a thread is issued before any LDS store, so they are in
// flight together.  A column past nzh loads a copy of column nzh - 1 (a valid
// address): columns transform independently and such a column is never
// stored back.
// The length-L twiddles load with them, into tw.
template <int L, int KB>
__device__ __forceinline__ void load_cols(const float2 *__restrict__ base, size_t lstride, int kz0, int nzh,
                                          float2 *buf, const float2 *__restrict__ tw_all, float2 *tw) {
  constexpr int total = L * KB, iters = (total + kFftThreads - 1) / kFftThreads;
  float2 v[iters];
#pragma unroll
  for (int it = 0; it < iters; ++it) {
    const int f0 = threadIdx.x + it * kFftThreads;
    const int f = (total % kFftThreads != 0 && f0 >= total) ? total - 1 : f0;
    const int c = f & (KB - 1), l = f / KB;
    v[it] = base[size_t(l) * lstride + (kz0 + c < nzh ? kz0 + c : nzh - 1)];
  }
  TwStage<L> tws;
  tws.load(tw_all, tw_off(L));
  tws.store(tw);
#pragma unroll
  for (int it = 0; it < iters; ++it) {
    const int f = threadIdx.x + it * kFftThreads;
    if (total % kFftThreads == 0 || f < total) buf[f] = v[it];
  }
}

// Column FFTs along one axis of a half spectrum, in place: column c of
// workgroup (blockIdx.x, outer = blockIdx.y, batch item blockIdx.z) is kz =
// blockIdx.x * KB + c, element l at item + outer * ostride + l * lstride + kz.
template <int L, bool INV>
__global__ void __launch_bounds__(kFftThreads)
fft_cols_kernel(float2 *__restrict__ data, int nzh, size_t lstride, size_t ostride, size_t bstride,
                const float2 *__restrict__ tw_all, uint32_t nkzb, uint32_t nouter, uint32_t nblk, uint32_t per) {
  constexpr int KB = FftCols<L>::KB, total = L * KB;
  __shared__ float2 buf[total];
  __shared__ float2 tw[L];
  // workgroup b -> block (b mod 8) per + b / 8: the dispatcher deals
  // workgroups to the 8 XCDs in turn, so each XCD walks one contiguous run of
  // blocks and the kz blocks of a row, which share cache lines at their
  // edges, run on one XCD
  const uint32_t item = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  if (item >= nblk) return;
  const uint32_t kzb = item % nkzb, rest = item / nkzb, outer = rest % nouter, bi = rest / nouter;
  float2 *base = data + size_t(bi) * bstride + size_t(outer) * ostride;
  const int kz0 = int(kzb) * KB;
  constexpr int iters = (total + kFftThreads - 1) / kFftThreads;
  load_cols<L, KB>(base, lstride, kz0, nzh, buf, tw_all, tw);
  __syncthreads();
  lds_fft<L, KB, KB, INV>(buf, tw);
#pragma unroll
  for (int it = 0; it < iters; ++it) {
    const int f = threadIdx.x + it * kFftThreads;
    if (total % kFftThreads != 0 && f >= total) break;
    const int c = f & (KB - 1), l = f / KB;
    if (kz0 + c < nzh) base[size_t(l) * lstride + kz0 + c] = buf[f];
  }
}

// The x pass of a potential solve, in place: the forward column FFT, the
// potential multiplier at each mode, the inverse column FFT.  Columns as in
// fft_cols_kernel with the x axis as the column axis (element l = kx) and
// outer = ky; the same float operations as the forward x pass,
// potential_kernel and the inverse x pass run one after another.
template <int L>
__global__ void __launch_bounds__(kFftThreads)
fft_cols_potential_kernel(float2 *__restrict__ data, int nzh, size_t lstride, size_t ostride, size_t bstride,
                          const float2 *__restrict__ tw_all, uint32_t nkzb, uint32_t nouter, uint32_t nblk,
                          uint32_t per, Tabs t, float r2, float scale_s, const float *__restrict__ scale_dev,
                          int deconv) {
  constexpr int KB = FftCols<L>::KB, total = L * KB;
  __shared__ float2 buf[total];
  __shared__ float2 tw[L];
  const uint32_t item = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  if (item >= nblk) return;
  const uint32_t kzb = item % nkzb, rest = item / nkzb, outer = rest % nouter, bi = rest / nouter;
  float2 *base = data + size_t(bi) * bstride + size_t(outer) * ostride;
  const int kz0 = int(kzb) * KB;
  constexpr int iters = (total + kFftThreads - 1) / kFftThreads;
  // Element by element: batching these loads as the other passes do raised
  // its registers past 64 and cut the resident waves from 8 to 6 per SIMD
  TwStage<L> tws;
  tws.load(tw_all, tw_off(L));
  tws.store(tw);
#pragma unroll
  for (int it = 0; it < iters; ++it) {
    const int f = threadIdx.x + it * kFftThreads;
    if (total % kFftThreads != 0 && f >= total) break;
    const int c = f & (KB - 1), l = f / KB;
    buf[f] = kz0 + c < nzh ? base[size_t(l) * lstride + kz0 + c] : make_float2(0.0f, 0.0f);
  }
  __syncthreads();
  lds_fft<L, KB, KB, false>(buf, tw);
  const float scale = scale_dev ? scale_dev[bi] : scale_s;
  const uint32_t jy = t.oy + outer;
#pragma unroll
  for (int it = 0; it < iters; ++it) {
    const int f = threadIdx.x + it * kFftThreads;
    if (total % kFftThreads != 0 && f >= total) break;
    const int c = f & (KB - 1), l = f / KB;
    if (kz0 + c < nzh) {
      const float fac = potential_factor(t, uint32_t(l), jy, t.oz + uint32_t(kz0 + c), r2, deconv, scale);
      const float2 v = buf[f];
      buf[f] = make_float2(v.x * fac, v.y * fac);
    }
  }
  __syncthreads();
  lds_fft<L, KB, KB, true>(buf, tw);
#pragma unroll
  for (int it = 0; it < iters; ++it) {
    const int f = threadIdx.x + it * kFftThreads;
    if (total % kFftThreads != 0 && f >= total) break;
    const int c = f & (KB - 1), l = f / KB;
    if (kz0 + c < nzh) base[size_t(l) * lstride + kz0 + c] = buf[f];
  }
}

// The three channels of a vector mesh whose divergence the R2C reads in
// place of a scalar mesh: channel d of item b at g3 + b * 2 * in_bstride +
// d * chan floats.
struct DivIn {
  const float *g3;
  size_t chan;
  uint32_t nx, ny;
  FastDiv div_ny;   // row -> (x, y)
};

// sum_d D_d g_d at cells (x, y, 2l) and (x, y, 2l + 1) of a mesh with
// nz = 2M: the fourth-order central difference (8 (g[+1] - g[-1]) -
// (g[+2] - g[-2])) / 12 along each axis, periodic, in the operation order
// of fd4_divergence_kernel with scale 1.  Each channel is read as float2
// pairs along z.
template <int M>
__device__ __forceinline__ float2 divergence_pair(const float *__restrict__ g, size_t chan, uint32_t nx, uint32_t ny,
                                                  uint32_t x, uint32_t y, int l) {
  constexpr float c8 = 8.0f / 12.0f, c1 = 1.0f / 12.0f;
  const float2 *gx = reinterpret_cast<const float2 *>(g);
  const float2 *gy = reinterpret_cast<const float2 *>(g + chan);
  const float2 *gz = reinterpret_cast<const float2 *>(g + 2 * chan);
  auto at = [&](const float2 *ch, uint32_t xi, uint32_t yi, int li) {
    return ch[(size_t(xi) * ny + yi) * M + uint32_t(li)];
  };
  // periodic neighbours (nx, ny >= 2, x < nx, y < ny): i + d wrapped by one
  // compare, d in {-2, -1, 1, 2}
  auto wrap = [](uint32_t i, int d, uint32_t n) {
    const uint32_t v = i + n + uint32_t(d);
    return v >= 2u * n ? v - 2u * n : (v >= n ? v - n : v);
  };
  const float2 xp1 = at(gx, wrap(x, 1, nx), y, l), xm1 = at(gx, wrap(x, -1, nx), y, l);
  const float2 xp2 = at(gx, wrap(x, 2, nx), y, l), xm2 = at(gx, wrap(x, -2, nx), y, l);
  const float2 yp1 = at(gy, x, wrap(y, 1, ny), l), ym1 = at(gy, x, wrap(y, -1, ny), l);
  const float2 yp2 = at(gy, x, wrap(y, 2, ny), l), ym2 = at(gy, x, wrap(y, -2, ny), l);
  // z - 2, z - 1 | z, z + 1 | z + 2, z + 3
  const float2 zl = at(gz, x, y, (l + M - 1) % M), zc = at(gz, x, y, l), zr = at(gz, x, y, (l + 1) % M);
  const float dx0 = c8 * (xp1.x - xm1.x) - c1 * (xp2.x - xm2.x), dx1 = c8 * (xp1.y - xm1.y) - c1 * (xp2.y - xm2.y);
  const float dy0 = c8 * (yp1.x - ym1.x) - c1 * (yp2.x - ym2.x), dy1 = c8 * (yp1.y - ym1.y) - c1 * (yp2.y - ym2.y);
  const float dz0 = c8 * (zc.y - zl.y) - c1 * (zr.x - zl.x), dz1 = c8 * (zr.x - zc.x) - c1 * (zr.y - zl.y);
  return make_float2((dx0 + dy0) + dz0, (dx1 + dy1) + dz1);
}

// z pass of the R2C: rows r = blockIdx.x * KB + c (< nrows) of item
// blockIdx.y, each nz = 2M reals read as M complex points.  With DIV the
// reals are the divergence of the vector mesh dv, formed as they load.
template <int M, bool DIV>
__global__ void __launch_bounds__(kFftThreads)
fft_r2c_rows_kernel(const float2 *__restrict__ in, float2 *__restrict__ out, int nrows, size_t in_bstride,
                    size_t out_bstride, const float2 *__restrict__ tw_all, DivIn dv, FaceCells fc) {
  constexpr int KB = FftCols<M>::KB, PITCH = KB + 1, total = M * KB;
  __shared__ float2 buf[M * PITCH];
  __shared__ float2 twM[M];
  __shared__ float2 twN[M + 1];
  // twM, and W_{2M}^k for the split, k = 0 .. M: loaded with the data and
  // ordered before their use by the barrier after it
  TwStage<M> tsM;
  TwStage<M + 1> tsN;
  const int r0 = blockIdx.x * KB;
  float2 *dst = out + size_t(blockIdx.y) * out_bstride;
  // rows are contiguous: lanes along l
  constexpr int iters = (total + kFftThreads - 1) / kFftThreads;
  if constexpr (DIV) {
    const float *g = dv.g3 + size_t(blockIdx.y) * 2 * in_bstride;
#pragma unroll
    for (int it = 0; it < iters; ++it) {
      const int f = threadIdx.x + it * kFftThreads;
      if (total % kFftThreads != 0 && f >= total) break;
      const int l = f % M, c = f / M;
      const uint32_t r = uint32_t(r0 + c);
      const uint32_t x = dv.div_ny.div(r);
      buf[l * PITCH + c] = r0 + c < nrows ? divergence_pair<M>(g, dv.chan, dv.nx, dv.ny, x, r - x * dv.ny, l)
                                          : make_float2(0.0f, 0.0f);
    }
    tsM.load(tw_all, tw_off(M));
    tsN.load(tw_all, tw_off(2 * M));
  } else {
    const float2 *src = in + size_t(blockIdx.y) * in_bstride;
    // all loads of a thread first, then the LDS stores; a row past nrows
    // loads a copy of row nrows - 1 (rows transform independently and such a
    // row is never stored)
    float2 v[iters];
    // with face cells apart: a pair in a tile's lx = 0 or ly = 0 plane loads
    // from the face buffer instead of the mesh; the lz = 0 cell of any other
    // pair (its x) is loaded from there too, by every lane (a partner lane's
    // line), and selected after all loads, so that no load waits on a branch
    float zf[iters];
    bool zsel[iters];
#pragma unroll
    for (int it = 0; it < iters; ++it) {
      const int f0 = threadIdx.x + it * kFftThreads;
      const int f = (total % kFftThreads != 0 && f0 >= total) ? total - 1 : f0;
      const int l = f % M, c = f / M;
      const uint32_t r = uint32_t(r0 + c < nrows ? r0 + c : nrows - 1);
      const float2 *p = src + size_t(r) * M + l;
      zsel[it] = false;
      if (fc.fb) {
        // cells (x, y, 2l) and (x, y, 2l + 1): tile (x/16, y/16, l/8), local
        // (a, b, cz) and (a, b, cz + 1) with cz = 2 (l mod 8); face offsets
        // in 32 bits (the host checks the buffer's size)
        const uint32_t x = fc.div_ny.div(r), y = r - x * fc.ny;
        const uint32_t a = x & uint32_t(kTile - 1), b = y & uint32_t(kTile - 1), cz = (uint32_t(l) & 7u) * 2u;
        const uint32_t tile = ((x >> 4) * fc.ty + (y >> 4)) * fc.tz + (uint32_t(l) >> 3);
        // face_index(a, b, 0): in a plane (a == 0 or b == 0) the row's first
        // cell, whose pair at cz is consecutive there; else the lz = 0 cell
        const uint32_t row = a == 0 ? b * uint32_t(kTile)
                                    : (b == 0 ? uint32_t(kTile * kTile) + (a - 1u) * uint32_t(kTile)
                                              : uint32_t(kTile * kTile + (kTile - 1) * kTile) +
                                                    (a - 1u) * uint32_t(kTile - 1) + (b - 1u));
        const float *ft = fc.fb + (tile * uint32_t(kFaceStride) + row);
        const bool plane = a == 0 || b == 0;
        if (plane) p = reinterpret_cast<const float2 *>(ft + cz);
        zf[it] = *ft;
        zsel[it] = !plane && cz == 0;
      }
      v[it] = *p;
    }
    tsM.load(tw_all, tw_off(M));
    tsN.load(tw_all, tw_off(2 * M));
#pragma unroll
    for (int it = 0; it < iters; ++it) {
      const int f = threadIdx.x + it * kFftThreads;
      if (total % kFftThreads != 0 && f >= total) break;
      const int l = f % M, c = f / M;
      if (zsel[it]) v[it].x = zf[it];
      buf[l * PITCH + c] = v[it];
    }
  }
  tsM.store(twM);
  tsN.store(twN);
  __syncthreads();
  lds_fft<M, KB, PITCH, false>(buf, twM);
  // X[k] = E[k] + W_{2M}^k O[k], E = (Z[k] + conj Z[M-k]) / 2,
  // O = (Z[k] - conj Z[M-k]) / 2i, k = 0 .. M
  constexpr int nzh = M + 1, tot2 = nzh * KB;
  for (int f = threadIdx.x; f < tot2; f += kFftThreads) {
    const int c = f / nzh, k = f - c * nzh;
    if (r0 + c >= nrows) continue;
    const float2 z = buf[(k % M) * PITCH + c];
    const float2 zm = buf[((M - k) % M) * PITCH + c];
    const float2 E = make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y));
    const float2 O = make_float2(0.5f * (z.y + zm.y), -0.5f * (z.x - zm.x));
    dst[size_t(r0 + c) * nzh + k] = cadd(E, cmul(twN[k], O));
  }
}

// z pass of the C2R: half-spectrum rows (M + 1 values) -> rows of 2M reals,
// in workgroups of kC2rThreads threads over up to kC2rPoints points (at 256^3
// 0.0263 -> 0.0245 ms against the 512-thread, 4096-point form the other
// passes use, which are slower in the small form).
constexpr int kC2rThreads = 256;
constexpr int kC2rPoints = 2048;
template <int M>
struct C2rRows {
  static constexpr int KB = pow2_floor((kC2rPoints / M) > 64 ? 64 : (kC2rPoints / M));
};

template <int M>
__global__ void __launch_bounds__(kC2rThreads)
fft_c2r_rows_kernel(const float2 *__restrict__ in, float2 *__restrict__ out, int nrows, size_t in_bstride,
                    size_t out_bstride, const float2 *__restrict__ tw_all) {
  constexpr int KB = C2rRows<M>::KB, PITCH = KB + 1, total = M * KB, nzh = M + 1;
  __shared__ float2 buf[M * PITCH];
  __shared__ float2 twM[M];
  TwStage<M, kC2rThreads> tsM;
  const float2 *twN = tw_all + tw_off(2 * M);
  const int r0 = blockIdx.x * KB;
  const float2 *src = in + size_t(blockIdx.y) * in_bstride;
  float2 *dst = out + size_t(blockIdx.y) * out_bstride;
  // Z'[k] = E' + i O', E' = X[k] + conj X[M-k], O' = (X[k] - conj X[M-k]) W_{2M}^{-k};
  // X[0] and X[M] enter with their real parts only.  Every load of a thread
  // is issued before the first use; a row past nrows loads a copy of row
  // nrows - 1 (rows transform independently and such a row is never stored).
  constexpr int iters = (total + kC2rThreads - 1) / kC2rThreads;
  float2 va[iters], vb[iters], vt[iters];
#pragma unroll
  for (int it = 0; it < iters; ++it) {
    const int f0 = threadIdx.x + it * kC2rThreads;
    const int f = (total % kC2rThreads != 0 && f0 >= total) ? total - 1 : f0;
    const int k = f % M, c = f / M;
    const float2 *row = src + size_t(r0 + c < nrows ? r0 + c : nrows - 1) * nzh;
    va[it] = row[k];
    vb[it] = row[M - k];
    vt[it] = twN[k];
  }
  tsM.load(tw_all, tw_off(M));
  tsM.store(twM);
#pragma unroll
  for (int it = 0; it < iters; ++it) {
    const int f = threadIdx.x + it * kC2rThreads;
    if (total % kC2rThreads != 0 && f >= total) break;
    const int k = f % M, c = f / M;
    float2 a = va[it], b = vb[it];
    if (k == 0) {
      a.y = 0.0f;
      b.y = 0.0f;
    }
    const float2 bc = conjf2(b);
    const float2 E = cadd(a, bc);
    const float2 O = cmul(csub(a, bc), conjf2(vt[it]));
    buf[k * PITCH + c] = make_float2(E.x - O.y, E.y + O.x);
  }
  __syncthreads();
  lds_fft<M, KB, PITCH, true, kC2rThreads>(buf, twM);
#pragma unroll
  for (int it = 0; it < iters; ++it) {
    const int f = threadIdx.x + it * kC2rThreads;
    if (total % kC2rThreads != 0 && f >= total) break;
    const int l = f % M, c = f / M;
    if (r0 + c < nrows) dst[size_t(r0 + c) * M + l] = buf[l * PITCH + c];
  }
}


// The native transforms cover meshes whose nx and ny are transform lengths
// (powers of two 2 .. 1024 or 3 * 2^k 6 .. 768; nx = 1 too: 2-D images, no
// x pass) and whose nz is twice one (the R2C's complex rows of nz / 2).
bool native_fft_ok(int nx, int ny, int nz) {
  return (nx == 1 || fft_len(nx)) && fft_len(ny) && nz % 2 == 0 && fft_len(nz / 2);
}

// Launch helpers over the compile-time lengths.
#define JPM_FFT_LENGTHS(X) \
  X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(6) X(12) X(24) X(48) X(96) X(192) X(384) X(768)

static int cols_pass(hipStream_t s, bool inv, float2 *data, int L, int nzh, int outer, size_t lstride,
                     size_t ostride, size_t bstride, int batch, const float2 *tw) {
  JPM_TIMED(s, "fft_cols_kernel");
  switch (L) {
#define JPM_COLS(N)                                                                                    \
  case N: {                                                                                            \
    const uint32_t nkzb = (nzh + FftCols<N>::KB - 1) / FftCols<N>::KB;                                \
    const uint32_t nblk = nkzb * uint32_t(outer) * uint32_t(batch), per = (nblk + 7) / 8;              \
    auto kern = inv ? fft_cols_kernel<N, true> : fft_cols_kernel<N, false>;                             \
    hipLaunchKernelGGL(kern, dim3(8 * per), dim3(kFftThreads), 0, s, data, nzh, lstride, ostride,      \
                       bstride, tw, nkzb, uint32_t(outer), nblk, per);                                 \
    break;                                                                                             \
  }
    JPM_FFT_LENGTHS(JPM_COLS)
#undef JPM_COLS
    default:
      return fail(JPM_E_INVALID, "native FFT length %d", L);
  }
  JPM_LAUNCHED("fft_cols_kernel");
  return JPM_OK;
}

// The R2C's z pass of in, or with dv.g3 set of the divergence of dv's
// vector mesh.
static int rows_r2c(hipStream_t s, int nx, int ny, int nz, int batch, const float *in, float2 *out,
                    const float2 *tw, const DivIn &dv = DivIn{nullptr, 0, 0, 0, FastDiv(1)},
                    const FaceCells &fc = FaceCells{nullptr, 0, 0, 0, 0, FastDiv(1)}) {
  JPM_REQUIRE(!fc.fb || (!dv.g3 && batch == 1 && nx % kTile == 0 && ny % kTile == 0 && nz % kTile == 0 &&
                         fc.ny == uint32_t(ny) && fc.ty == uint32_t(ny / kTile) && fc.tz == uint32_t(nz / kTile) &&
                         size_t(nx / kTile) * size_t(ny / kTile) * size_t(nz / kTile) * kFaceStride <
                             (size_t(1) << 32)),
              "face cells apart need one mesh of full 16^3 tiles");
  const int M = nz / 2, nzh = M + 1, nrows = nx * ny;
  const size_t H = size_t(nx) * ny * nzh, Mc = size_t(nx) * ny * M;
  const float2 *i2 = reinterpret_cast<const float2 *>(in);
  JPM_TIMED(s, dv.g3 ? "fft_r2c_rows_kernel[div]" : "fft_r2c_rows_kernel");
  switch (M) {
#define JPM_R2C(N)                                                                                            \
  case N: {                                                                                                   \
    auto kern = dv.g3 ? fft_r2c_rows_kernel<N, true> : fft_r2c_rows_kernel<N, false>;                         \
    hipLaunchKernelGGL(kern, dim3((nrows + FftCols<N>::KB - 1) / FftCols<N>::KB, batch), dim3(kFftThreads), 0,  \
                       s, i2, out, nrows, Mc, H, tw, dv, fc);                                                 \
    break;                                                                                                    \
  }
    JPM_FFT_LENGTHS(JPM_R2C)
#undef JPM_R2C
    default:
      return fail(JPM_E_INVALID, "native FFT length %d", nz);
  }
  JPM_LAUNCHED("fft_r2c_rows_kernel");
  return JPM_OK;
}

static int rows_c2r(hipStream_t s, int nx, int ny, int nz, int batch, float2 *in, float *out, const float2 *tw) {
  const int M = nz / 2, nzh = M + 1, nrows = nx * ny;
  const size_t H = size_t(nx) * ny * nzh, Mc = size_t(nx) * ny * M;
  float2 *o2 = reinterpret_cast<float2 *>(out);
  JPM_TIMED(s, "fft_c2r_rows_kernel");
  switch (M) {
#define JPM_C2R(N)                                                                                            \
  case N:                                                                                                     \
    hipLaunchKernelGGL(fft_c2r_rows_kernel<N>, dim3((nrows + C2rRows<N>::KB - 1) / C2rRows<N>::KB, batch),    \
                       dim3(kC2rThreads), 0, s, in, o2, nrows, H, Mc, tw);                                    \
    break;
    JPM_FFT_LENGTHS(JPM_C2R)
#undef JPM_C2R
    default:
      return fail(JPM_E_INVALID, "native FFT length %d", nz);
  }
  JPM_LAUNCHED("fft_c2r_rows_kernel");
  return JPM_OK;
}

int native_r2c(hipStream_t s, int nx, int ny, int nz, int batch, const float *in, void *out) {
  JPM_REQUIRE(native_fft_ok(nx, ny, nz) && batch >= 1 && batch <= 65535, "native FFT shape");
  const float2 *tw;
  int rc = twiddles(&tw);
  if (rc) return rc;
  const int nzh = nz / 2 + 1;
  const size_t H = size_t(nx) * ny * nzh;
  float2 *o = static_cast<float2 *>(out);
  if ((rc = rows_r2c(s, nx, ny, nz, batch, in, o, tw))) return rc;
  if ((rc = cols_pass(s, false, o, ny, nzh, nx, size_t(nzh), size_t(ny) * nzh, H, batch, tw))) return rc;
  if (nx == 1) return JPM_OK;
  return cols_pass(s, false, o, nx, nzh, ny, size_t(ny) * nzh, size_t(nzh), H, batch, tw);
}

int native_c2r(hipStream_t s, int nx, int ny, int nz, int batch, void *in, float *out) {
  JPM_REQUIRE(native_fft_ok(nx, ny, nz) && batch >= 1 && batch <= 65535, "native FFT shape");
  const float2 *tw;
  int rc = twiddles(&tw);
  if (rc) return rc;
  const int nzh = nz / 2 + 1;
  const size_t H = size_t(nx) * ny * nzh;
  float2 *h = static_cast<float2 *>(in);
  if (nx > 1 && (rc = cols_pass(s, true, h, nx, nzh, ny, size_t(ny) * nzh, size_t(nzh), H, batch, tw))) return rc;
  if ((rc = cols_pass(s, true, h, ny, nzh, nx, size_t(nzh), size_t(ny) * nzh, H, batch, tw))) return rc;
  return rows_c2r(s, nx, ny, nz, batch, h, out, tw);
}

// out = C2R(scale G . R2C(in)) for a batch of meshes, the spectrum in spec
// (H complex per mesh): R2C rows and y columns, the x columns with the
// multiplier between their forward and inverse transforms, inverse y
// columns, C2R rows.  in and out may be the same array.  With g3 the input
// is instead sum_d D_d g_d of the vector meshes at g3 (channels chan floats
// apart, items nx ny nz floats apart), formed by the R2C rows as they load.
int native_potential(hipStream_t s, int nx, int ny, int nz, int batch, const float *in, void *spec, float *out,
                     const Tabs &t, float r2, float scale, const float *scale_dev, int deconv, const float *g3,
                     size_t chan, const FaceCells *faces) {
  JPM_REQUIRE(native_fft_ok(nx, ny, nz) && nx >= 2 && batch >= 1 && batch <= 65535, "native FFT shape");
  const float2 *tw;
  int rc = twiddles(&tw);
  if (rc) return rc;
  const int nzh = nz / 2 + 1;
  const size_t H = size_t(nx) * ny * nzh;
  float2 *h = static_cast<float2 *>(spec);
  DivIn dv{g3, chan, uint32_t(nx), uint32_t(ny), FastDiv(uint32_t(ny))};
  const FaceCells fc = faces ? *faces : FaceCells{nullptr, 0, 0, 0, 0, FastDiv(1)};
  if ((rc = rows_r2c(s, nx, ny, nz, batch, in, h, tw, dv, fc))) return rc;
  if ((rc = cols_pass(s, false, h, ny, nzh, nx, size_t(nzh), size_t(ny) * nzh, H, batch, tw))) return rc;
  {
    JPM_TIMED(s, "fft_cols_potential_kernel");
    switch (nx) {
#define JPM_POT(N)                                                                                        \
    case N: {                                                                                             \
      const uint32_t nkzb = (nzh + FftCols<N>::KB - 1) / FftCols<N>::KB;                                 \
      const uint32_t nblk = nkzb * uint32_t(ny) * uint32_t(batch), per = (nblk + 7) / 8;                  \
      hipLaunchKernelGGL(fft_cols_potential_kernel<N>, dim3(8 * per), dim3(kFftThreads), 0, s, h, nzh,    \
                         size_t(ny) * nzh, size_t(nzh), H, tw, nkzb, uint32_t(ny), nblk, per, t, r2,      \
                         scale, scale_dev, deconv);                                                       \
      break;                                                                                              \
    }
      JPM_FFT_LENGTHS(JPM_POT)
#undef JPM_POT
      default:
        return fail(JPM_E_INVALID, "native FFT length %d", nx);
    }
  }
  JPM_LAUNCHED("fft_cols_potential_kernel");
  if ((rc = cols_pass(s, true, h, ny, nzh, nx, size_t(nzh), size_t(ny) * nzh, H, batch, tw))) return rc;
  return rows_c2r(s, nx, ny, nz, batch, h, out, tw);
}

}  // namespace jpm
