// Shared host/device helpers for libjaxpm_hip.so (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "jaxpm_hip.h"

namespace jpm {

// Records a formatted message as the calling thread's last error and
// returns `code`, so entry points can `return fail(JPM_E_INVALID, ...)`.
int fail(int code, const char *fmt, ...);
void clear_error();

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

#define JPM_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return ::jpm::fail(JPM_E_HIP, "%s failed: %s (%s:%d)", #call,            \
                         hipGetErrorString(e_), __FILE__, __LINE__);           \
  } while (0)

#define JPM_REQUIRE(cond, ...)                                                 \
  do {                                                                         \
    if (!(cond)) return ::jpm::fail(JPM_E_INVALID, __VA_ARGS__);               \
  } while (0)

#define JPM_LAUNCHED(name)                                                     \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess)                                                      \
      return ::jpm::fail(JPM_E_HIP, "launch of %s failed: %s", name,           \
                         hipGetErrorString(e_));                               \
  } while (0)

// Optional per-kernel timing: when enabled (jpm_timing_enable), launches
// wrapped in JPM_TIMED record a pair of HIP events on their stream; the
// report sums the elapsed times per kernel name.
bool timing_enabled();
// While one is alive, the calling thread records no timing events.
struct TimingPause {
  TimingPause();
  ~TimingPause();
  TimingPause(const TimingPause &) = delete;
  TimingPause &operator=(const TimingPause &) = delete;
};
void timing_record(hipStream_t s, const char *name, hipEvent_t begin, hipEvent_t end);

struct TimedScope {
  hipStream_t s;
  const char *name;
  hipEvent_t a = nullptr, b = nullptr;
  TimedScope(hipStream_t st, const char *nm) : s(st), name(nm) {
    if (timing_enabled() && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
      (void)hipEventRecord(a, s);
    else
      a = b = nullptr;
  }
  ~TimedScope() {
    if (a && b) {
      (void)hipEventRecord(b, s);
      timing_record(s, name, a, b);
    }
  }
};

#define JPM_TIMED(stream, name) ::jpm::TimedScope jpm_timed_scope_(stream, name)

// Largest mesh the 32-bit flat cell index supports (1290^3 cells).
constexpr int64_t kMaxCells = (int64_t(1) << 31) - 1;

inline int check_dims(int nx, int ny, int nz) {
  JPM_REQUIRE(nx > 0 && ny > 0 && nz > 0, "mesh shape (%d,%d,%d) must be positive", nx, ny, nz);
  JPM_REQUIRE(int64_t(nx) * ny * nz <= kMaxCells, "mesh of %lld cells exceeds the 2^31-1 limit",
              (long long)(int64_t(nx) * ny * nz));
  return JPM_OK;
}

inline int64_t cells(int nx, int ny, int nz) { return int64_t(nx) * ny * nz; }
inline int64_t half_cells(int nx, int ny, int nz) { return int64_t(nx) * ny * (nz / 2 + 1); }

// Unsigned division by a divisor fixed at launch time with one multiply-high
// (Granlund & Montgomery 1994, fig. 4.1); exact for all 32-bit numerators and
// divisors 1 <= d < 2^31.
struct FastDiv {
  uint32_t d, m, sh1, sh2;
  FastDiv() = default;
  explicit FastDiv(uint32_t div) : d(div) {
    uint32_t l = 0;
    while ((uint64_t(1) << l) < div) ++l;
    m = uint32_t(((uint64_t(1) << 32) * ((uint64_t(1) << l) - div)) / div + 1);
    sh1 = l < 1 ? l : 1;
    sh2 = l > 1 ? l - 1 : 0;
  }
  __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const {
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t t = __umulhi(n, m);
#else
    uint32_t t = uint32_t((uint64_t(n) * m) >> 32);
#endif
    return (t + ((n - t) >> sh1)) >> sh2;
  }
};

inline size_t align_up(size_t x, size_t a = 256) { return (x + a - 1) / a * a; }

// int32(f) of a floor value f, with NaN, +-Inf and values outside the int32
// range mapped to 0 (a C++ float-to-int conversion of those is undefined):
// non-finite positions then index valid cells, with NaN weights.
__host__ __device__ inline int cell_int(float f) {
  return (f >= -2147483648.0f && f < 2147483648.0f) ? int(f) : 0;
}

// |x| for finite x and 0 for NaN or +-Inf: the bound of the fixed-point
// scale covers the finite deposits only (non-finite ones bypass the
// fixed-point sums and reach the mesh as floats, so NaN and Inf propagate to
// the cells they touch as in a float scatter-add).
__host__ __device__ inline float finite_abs(float x) {
  const float a = fabsf(x);
  return a < INFINITY ? a : 0.0f;
}

// 2^e with e the largest exponent such that n * wmax * 2^e < 2^62 and
// wmax * 2^e < 2^51: the scale of the 64-bit fixed-point sums of the
// deterministic paint.  Every particle spreads |w| over its corners (the CIC
// weights sum to 1), so no cell of the fixed-point sum can leave the int64
// range, and every single deposit stays below 2^51 in magnitude, the range
// fixed_point_q converts exactly (it binds only below 2^11 particles).
__host__ __device__ inline double fixed_point_scale(float wmax, uint32_t n) {
  double bound = double(wmax) * double(n > 0 ? n : 1);
  if (!(bound > 0.0) || !(bound < 1e300)) return 1.0;
  int e, ew;
  frexp(bound, &e);          // bound < 2^e
  frexp(double(wmax), &ew);  // wmax < 2^ew
  return ldexp(1.0, (62 - e) < (51 - ew) ? (62 - e) : (51 - ew));
}

// rint(double(v) * scale) as a 64-bit two's complement integer, for a power
// of two `scale` and |v| * scale < 2^51.  The product is exact, and adding
// 1.5 * 2^52 in the same fused operation rounds it to the nearest integer
// (ties to even, as rint) inside the binade [2^52, 2^53), whose doubles are
// the consecutive integers: the result's bit pattern minus that of 1.5 * 2^52
// (0x4338000000000000, low word zero) is the integer itself.  One fma and one
// 32-bit subtract instead of a rounding and a double-to-int64 conversion.
__device__ __forceinline__ unsigned long long fixed_point_q(float v, double scale) {
  const double t = fma(double(v), scale, 6755399441055744.0);
  const uint32_t lo = uint32_t(__double2loint(t));
  const uint32_t hi = uint32_t(__double2hiint(t)) - 0x43380000u;
  return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// Number of blocks for a grid-stride loop over `n` items, capped so every
// XCD gets many workgroups without launching more than the loop needs.
inline unsigned grid_for(int64_t n, int block, int64_t cap = 256 * 64) {
  int64_t g = (n + block - 1) / block;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return unsigned(g);
}

}  // namespace jpm
