// Binned power spectra over R2C half spectra: the mode sums behind jaxpm
// utils.py:76-128 (power_spectrum) and its bin table utils.py:14-73.
//
// Each wave walks one (kx, ky) row of the half spectrum, lanes on
// consecutive kz (coalesced).  Per mode it forms |k| exactly as the
// reference bins it (float64 sum of squares, rounded to float32, float32
// square root), finds the bin by binary search over the float64 edges, and
// adds the mode and its Hermitian mirror (X(-k) = conj X(k), not stored in
// the half spectrum) to per-bin sums.  Lanes of one bin are summed across
// the wave first (one LDS float64 atomic per bin run), each workgroup writes
// its partial sums, and a second kernel adds the partials in a fixed order,
// so results repeat bitwise.
#include <cmath>

#include "common.hpp"

namespace jpm {

constexpr int kPkMaxPoles = 8;
constexpr int kPkMaxSlots = 8192;   // accumulators x bins, float64 in LDS (64 KiB)
constexpr int kPkBlock = 256;
constexpr int kPkMaxBlocks = 512;

struct PkParams {
  int nx, ny, nz, nzh;
  double fac[3];   // 2 pi n_d / L_d: cell wave number to h/Mpc
  double los[3];   // unit line of sight (multipoles only)
  int nedges, nbins, nslots, npoles;
  int cross, aniso;
  int poles[kPkMaxPoles];
};

// numpy.fft.fftfreq(n)[i] = i' * (1.0 / n), i' the signed index.
__device__ __forceinline__ double fftfreq(int i, int n) {
  const int s = i <= (n - 1) / 2 ? i : i - n;
  return double(s) * (1.0 / double(n));
}

// Correctly rounded float32 square root: a float64 root rounded to float32,
// then moved to the neighbour whose rounding interval holds f.  The interval
// ends (midpoints between adjacent floats, 25 significant bits) square
// exactly in float64.
__device__ __forceinline__ float sqrt_rn(float f) {
  if (!(f > 0.0f)) return 0.0f;
  float s = float(sqrt(double(f)));
  const double hi = 0.5 * (double(s) + double(nextafterf(s, INFINITY)));
  if (double(f) > hi * hi) return nextafterf(s, INFINITY);
  const double lo = 0.5 * (double(s) + double(nextafterf(s, 0.0f)));
  if (double(f) < lo * lo) return nextafterf(s, 0.0f);
  return s;
}

// Legendre polynomial by Bonnet's recurrence.
__device__ __forceinline__ double legendre(int ell, double mu) {
  if (ell == 0) return 1.0;
  double p0 = 1.0, p1 = mu;
  for (int n = 1; n < ell; ++n) {
    const double p2 = (double(2 * n + 1) * mu * p1 - double(n) * p0) / double(n + 1);
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

// np.digitize(x, edges) for increasing edges: the number of edges <= x.
__device__ __forceinline__ int digitize(const double *__restrict__ edges, int n, double x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (edges[mid] <= x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Slots per bin: 0 mode count, 1 sum of |k|, 2+q Re sum for pole q, and for
// cross spectra 2+npoles+q the Im sum.  acc is [slot][bin].
__global__ __launch_bounds__(kPkBlock) void pk_bin_kernel(const float2 *__restrict__ s1,
                                                          const float2 *__restrict__ s2,
                                                          const double *__restrict__ edges, PkParams p,
                                                          double *__restrict__ partial) {
  extern __shared__ double acc[];
  for (int e = threadIdx.x; e < p.nslots; e += blockDim.x) acc[e] = 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int waves = blockDim.x >> 6;
  const int64_t rows = int64_t(p.nx) * p.ny;
  for (int64_t row = int64_t(blockIdx.x) * waves + wave; row < rows; row += int64_t(gridDim.x) * waves) {
    const int i = int(row / p.ny);
    const int j = int(row - int64_t(i) * p.ny);
    const double kx = p.fac[0] * fftfreq(i, p.nx);
    const double ky = p.fac[1] * fftfreq(j, p.ny);
    // the Hermitian mirror's components (Nyquist entries mirror onto themselves)
    const double kxm = p.fac[0] * fftfreq(i == 0 ? 0 : p.nx - i, p.nx);
    const double kym = p.fac[1] * fftfreq(j == 0 ? 0 : p.ny - j, p.ny);
    const float2 *r1 = s1 + row * p.nzh;
    const float2 *r2 = s2 + row * p.nzh;
    for (int k0 = 0; k0 < p.nzh; k0 += 64) {
      const int k = k0 + lane;
      const bool valid = k < p.nzh;
      int bin = 0;
      double cnt = 0.0, ks = 0.0;
      double re[kPkMaxPoles], im[kPkMaxPoles];
#pragma unroll
      for (int q = 0; q < kPkMaxPoles; ++q) re[q] = im[q] = 0.0;
      if (valid) {
        const double kz = p.fac[2] * fftfreq(k, p.nz);
        const double ksq = kx * kx + ky * ky + kz * kz;
        const float kf = sqrt_rn(float(ksq));
        bin = digitize(edges, p.nedges, double(kf));
        const bool pair = k > 0 && 2 * k != p.nz;   // mirror not stored
        const float2 a = r1[k];
        const float2 b = p.cross ? r2[k] : a;
        const double xr = double(a.x) * double(b.x) + double(a.y) * double(b.y);
        const double xi = double(a.y) * double(b.x) - double(a.x) * double(b.y);
        cnt = pair ? 2.0 : 1.0;
        ks = cnt * double(kf);
        double mu = 1.0, mum = 1.0;
        if (p.aniso) {
          const double kzm = p.fac[2] * fftfreq(k == 0 ? 0 : p.nz - k, p.nz);
          const double num = kx * p.los[0] + ky * p.los[1] + kz * p.los[2];
          const double numm = kxm * p.los[0] + kym * p.los[1] + kzm * p.los[2];
          mu = kf == 0.0f ? 0.0 : num / double(kf);
          mum = kf == 0.0f ? 0.0 : numm / double(kf);
        }
#pragma unroll
        for (int q = 0; q < kPkMaxPoles; ++q) {
          if (q < p.npoles) {
            const int ell = p.poles[q];
            const double w = double(2 * ell + 1) * legendre(ell, mu);
            const double wm = pair ? double(2 * ell + 1) * legendre(ell, mum) : 0.0;
            re[q] = (w + wm) * xr;   // the mirror carries conj: Re adds, Im subtracts
            im[q] = (w - wm) * xi;
          }
        }
      }
      uint64_t todo = __ballot(valid);
      while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const int lb = __shfl(bin, leader);
        const bool mine = valid && bin == lb;
        todo &= ~uint64_t(__ballot(mine));
        double v = wave_sum(mine ? cnt : 0.0);
        if (lane == leader) atomicAdd(&acc[lb], v);
        v = wave_sum(mine ? ks : 0.0);
        if (lane == leader) atomicAdd(&acc[p.nbins + lb], v);
#pragma unroll
        for (int q = 0; q < kPkMaxPoles; ++q) {
          if (q < p.npoles) {
            v = wave_sum(mine ? re[q] : 0.0);
            if (lane == leader) atomicAdd(&acc[(2 + q) * p.nbins + lb], v);
            if (p.cross) {
              v = wave_sum(mine ? im[q] : 0.0);
              if (lane == leader) atomicAdd(&acc[(2 + p.npoles + q) * p.nbins + lb], v);
            }
          }
        }
      }
    }
  }
  __syncthreads();
  double *mine = partial + int64_t(blockIdx.x) * p.nslots;
  for (int e = threadIdx.x; e < p.nslots; e += blockDim.x) mine[e] = acc[e];
}

__global__ void pk_reduce_kernel(const double *__restrict__ partial, int blocks, int nslots,
                                 double *__restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nslots) return;
  double s = 0.0;
  for (int b = 0; b < blocks; ++b) s += partial[int64_t(b) * nslots + e];
  out[e] = s;
}

static int pk_blocks(int nx, int ny) {
  const int64_t rows = int64_t(nx) * ny;
  int64_t b = (rows + (kPkBlock / 64) - 1) / (kPkBlock / 64);
  return int(b < kPkMaxBlocks ? (b < 1 ? 1 : b) : kPkMaxBlocks);
}

int pk_slots(int nedges, int npoles, int cross) { return (2 + npoles * (cross ? 2 : 1)) * (nedges + 1); }

size_t power_spectrum_ws(int nx, int ny, int nz, int nedges, int npoles, int cross) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || nedges < 0 || npoles < 1) return 0;
  return size_t(pk_blocks(nx, ny)) * size_t(pk_slots(nedges, npoles, cross)) * sizeof(double);
}

int power_spectrum(hipStream_t s, int nx, int ny, int nz, const void *spec1, const void *spec2, double fx, double fy,
                   double fz, const double *edges, int nedges, const int *poles, int npoles, double lx, double ly,
                   double lz, double *out, void *ws, size_t ws_bytes) {
  if (int rc = check_dims(nx, ny, nz)) return rc;
  JPM_REQUIRE(spec1 && out, "power spectrum: null spectrum or output");
  JPM_REQUIRE(nedges >= 0 && (nedges == 0 || edges), "power spectrum: %d edges", nedges);
  JPM_REQUIRE(npoles >= 1 && npoles <= kPkMaxPoles && poles, "power spectrum: 1..%d multipoles, got %d",
              kPkMaxPoles, npoles);
  const int cross = spec2 != nullptr;
  const int nslots = pk_slots(nedges, npoles, cross);
  JPM_REQUIRE(nslots <= kPkMaxSlots, "power spectrum: %d bins x %d sums exceed %d", nedges + 1,
              nslots / (nedges + 1), kPkMaxSlots);
  JPM_REQUIRE(ws_bytes >= power_spectrum_ws(nx, ny, nz, nedges, npoles, cross) && ws,
              "power spectrum: workspace of %zu bytes, need %zu", ws_bytes,
              power_spectrum_ws(nx, ny, nz, nedges, npoles, cross));
  PkParams p{};
  p.nx = nx;
  p.ny = ny;
  p.nz = nz;
  p.nzh = nz / 2 + 1;
  p.fac[0] = fx;
  p.fac[1] = fy;
  p.fac[2] = fz;
  p.los[0] = lx;
  p.los[1] = ly;
  p.los[2] = lz;
  p.nedges = nedges;
  p.nbins = nedges + 1;
  p.nslots = nslots;
  p.npoles = npoles;
  p.cross = cross;
  p.aniso = 0;
  for (int q = 0; q < npoles; ++q) {
    JPM_REQUIRE(poles[q] >= 0 && poles[q] <= 64, "power spectrum: multipole %d outside [0, 64]", poles[q]);
    p.poles[q] = poles[q];
    if (poles[q] != 0) p.aniso = 1;
  }
  const int blocks = pk_blocks(nx, ny);
  double *partial = static_cast<double *>(ws);
  {
    JPM_TIMED(s, "pk_bin_kernel");
    hipLaunchKernelGGL(pk_bin_kernel, dim3(blocks), dim3(kPkBlock), size_t(nslots) * sizeof(double), s,
                       static_cast<const float2 *>(spec1), static_cast<const float2 *>(cross ? spec2 : spec1),
                       edges, p, partial);
    JPM_LAUNCHED("pk_bin_kernel");
  }
  {
    JPM_TIMED(s, "pk_reduce_kernel");
    hipLaunchKernelGGL(pk_reduce_kernel, dim3((nslots + 255) / 256), dim3(256), 0, s, partial, blocks, nslots, out);
    JPM_LAUNCHED("pk_reduce_kernel");
  }
  return JPM_OK;
}

}  // namespace jpm
