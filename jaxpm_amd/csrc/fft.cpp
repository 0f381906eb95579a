// hipFFT plan cache: one R2C / C2R plan per (shape, type, batch),
// created on first use and kept until jpm_shutdown.  Plans are bound to the
// caller's stream under the cache mutex right before execution, so two
// threads never race on one plan's stream.
#include <hipfft/hipfft.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <tuple>

#include "common.hpp"

namespace jpm {

namespace {
std::mutex g_mu;
std::map<std::tuple<int, int, int, int, int>, hipfftHandle> g_plans;

const char *fft_err(hipfftResult r) {
  switch (r) {
    case HIPFFT_SUCCESS: return "success";
    case HIPFFT_INVALID_PLAN: return "invalid plan";
    case HIPFFT_ALLOC_FAILED: return "allocation failed";
    case HIPFFT_INVALID_TYPE: return "invalid type";
    case HIPFFT_INVALID_VALUE: return "invalid value";
    case HIPFFT_INTERNAL_ERROR: return "internal error";
    case HIPFFT_EXEC_FAILED: return "exec failed";
    case HIPFFT_SETUP_FAILED: return "setup failed";
    case HIPFFT_INVALID_SIZE: return "invalid size";
    default: return "hipFFT error";
  }
}

// Caller holds g_mu.
int plan_locked(int nx, int ny, int nz, hipfftType type, int batch, hipfftHandle &out) {
  auto key = std::make_tuple(nx, ny, nz, int(type), batch);
  auto it = g_plans.find(key);
  if (it != g_plans.end()) {
    out = it->second;
    return JPM_OK;
  }
  hipfftHandle h;
  hipfftResult r = hipfftCreate(&h);
  if (r != HIPFFT_SUCCESS) return fail(JPM_E_FFT, "hipfftCreate: %s", fft_err(r));
  int n[3] = {nx, ny, nz};
  const long long M = cells(nx, ny, nz), H = half_cells(nx, ny, nz);
  const bool half_in = type == HIPFFT_C2R || type == HIPFFT_Z2D, half_out = type == HIPFFT_R2C || type == HIPFFT_D2Z;
  const int idist = int(half_in ? H : M);
  const int odist = int(half_out ? H : M);
  size_t work = 0;
  r = hipfftMakePlanMany(h, 3, n, nullptr, 1, idist, nullptr, 1, odist, type, batch, &work);
  if (r != HIPFFT_SUCCESS) {
    hipfftDestroy(h);
    return fail(JPM_E_FFT, "hipfftMakePlanMany(%d,%d,%d, batch %d): %s", nx, ny, nz, batch, fft_err(r));
  }
  // Plan creation may initialise device data (twiddle tables) on streams of
  // its own; finish that before the plan's first execution on the caller's
  // stream.  Plans are created once per shape, outside any captured graph.
  const hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    hipfftDestroy(h);
    return fail(JPM_E_HIP, "synchronising after FFT plan creation: %s", hipGetErrorString(e));
  }
  g_plans.emplace(key, h);
  out = h;
  return JPM_OK;
}
}  // namespace

// Batches execute as sub-batches of at most kMaxFftElems points each, so no
// plan's work area or offsets grow past what one 1024^3 transform needs.
constexpr long long kMaxFftElems = 1LL << 30;

static int sub_batch(int nx, int ny, int nz, int batch) {
  const long long m = cells(nx, ny, nz);
  long long sub = kMaxFftElems / m;
  if (sub < 1) sub = 1;
  return int(sub < batch ? sub : batch);
}

bool native_fft_ok(int nx, int ny, int nz);
int native_r2c(hipStream_t s, int nx, int ny, int nz, int batch, const float *in, void *out);
int native_c2r(hipStream_t s, int nx, int ny, int nz, int batch, void *in, float *out);

// Power-of-two meshes use the library's own transforms (fftk.hip); other
// shapes, or real buffers not 8-byte aligned, go to hipFFT.
static bool use_native(int nx, int ny, int nz, const void *real) {
  return native_fft_ok(nx, ny, nz) && (reinterpret_cast<uintptr_t>(real) & 7u) == 0;
}

int fft_r2c(hipStream_t s, int nx, int ny, int nz, int batch, const float *in, void *out) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(batch >= 1 && batch <= 65535, "batch %d out of range", batch);
  JPM_REQUIRE(in && out, "null FFT buffer");
  if (use_native(nx, ny, nz, in)) {
    JPM_TIMED(s, "native_r2c (3 passes)");
    return native_r2c(s, nx, ny, nz, batch, in, out);
  }
  const size_t M = size_t(cells(nx, ny, nz)), H = size_t(half_cells(nx, ny, nz));
  const int sub = sub_batch(nx, ny, nz, batch);
  std::lock_guard<std::mutex> lk(g_mu);
  JPM_TIMED(s, "hipfft_r2c");
  for (int b0 = 0; b0 < batch; b0 += sub) {
    const int nb = batch - b0 < sub ? batch - b0 : sub;
    hipfftHandle h;
    rc = plan_locked(nx, ny, nz, HIPFFT_R2C, nb, h);
    if (rc) return rc;
    hipfftResult r = hipfftSetStream(h, s);
    if (r == HIPFFT_SUCCESS)
      r = hipfftExecR2C(h, const_cast<hipfftReal *>(in) + size_t(b0) * M,
                        static_cast<hipfftComplex *>(out) + size_t(b0) * H);
    if (r != HIPFFT_SUCCESS) return fail(JPM_E_FFT, "hipfftExecR2C: %s", fft_err(r));
  }
  return JPM_OK;
}

int fft_c2r(hipStream_t s, int nx, int ny, int nz, int batch, void *in, float *out) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(batch >= 1 && batch <= 65535, "batch %d out of range", batch);
  JPM_REQUIRE(in && out, "null FFT buffer");
  if (use_native(nx, ny, nz, out)) {
    JPM_TIMED(s, "native_c2r (3 passes)");
    return native_c2r(s, nx, ny, nz, batch, in, out);
  }
  const size_t M = size_t(cells(nx, ny, nz)), H = size_t(half_cells(nx, ny, nz));
  const int sub = sub_batch(nx, ny, nz, batch);
  std::lock_guard<std::mutex> lk(g_mu);
  JPM_TIMED(s, "hipfft_c2r");
  for (int b0 = 0; b0 < batch; b0 += sub) {
    const int nb = batch - b0 < sub ? batch - b0 : sub;
    hipfftHandle h;
    rc = plan_locked(nx, ny, nz, HIPFFT_C2R, nb, h);
    if (rc) return rc;
    hipfftResult r = hipfftSetStream(h, s);
    if (r == HIPFFT_SUCCESS)
      r = hipfftExecC2R(h, static_cast<hipfftComplex *>(in) + size_t(b0) * H, out + size_t(b0) * M);
    if (r != HIPFFT_SUCCESS) return fail(JPM_E_FFT, "hipfftExecC2R: %s", fft_err(r));
  }
  return JPM_OK;
}

// Float64 transforms (hipFFT D2Z / Z2D) of a batch of meshes: the float64
// pm_forces.  Only these real-to-complex and complex-to-real plans are made.
int fft_d2z(hipStream_t s, int nx, int ny, int nz, int batch, const double *in, void *out) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(batch >= 1 && batch <= 3, "batch %d out of range", batch);
  JPM_REQUIRE(in && out, "null FFT buffer");
  std::lock_guard<std::mutex> lk(g_mu);
  hipfftHandle h;
  if ((rc = plan_locked(nx, ny, nz, HIPFFT_D2Z, batch, h))) return rc;
  hipfftResult r = hipfftSetStream(h, s);
  if (r == HIPFFT_SUCCESS) r = hipfftExecD2Z(h, const_cast<hipfftDoubleReal *>(in), static_cast<hipfftDoubleComplex *>(out));
  if (r != HIPFFT_SUCCESS) return fail(JPM_E_FFT, "hipfftExecD2Z: %s", fft_err(r));
  return JPM_OK;
}

int fft_z2d(hipStream_t s, int nx, int ny, int nz, int batch, void *in, double *out) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(batch >= 1 && batch <= 3, "batch %d out of range", batch);
  JPM_REQUIRE(in && out, "null FFT buffer");
  std::lock_guard<std::mutex> lk(g_mu);
  hipfftHandle h;
  if ((rc = plan_locked(nx, ny, nz, HIPFFT_Z2D, batch, h))) return rc;
  hipfftResult r = hipfftSetStream(h, s);
  if (r == HIPFFT_SUCCESS) r = hipfftExecZ2D(h, static_cast<hipfftDoubleComplex *>(in), out);
  if (r != HIPFFT_SUCCESS) return fail(JPM_E_FFT, "hipfftExecZ2D: %s", fft_err(r));
  return JPM_OK;
}

int c2c_split(hipStream_t s, int64_t m, int batch, const void *z, float *xy, float sign);
int c2c_combine(hipStream_t s, int nx, int ny, int nz, int batch, const void *xyh, void *out, float sign);

// Workspace of fft_c2c: the real and imaginary parts of every array as real
// fields [2 batch][M], then their half spectra [2 batch][H].
size_t fft_c2c_workspace_size(int nx, int ny, int nz, int batch) {
  const size_t M = size_t(cells(nx, ny, nz)), H = size_t(half_cells(nx, ny, nz));
  return align_up(2 * size_t(batch) * M * 4) + 2 * size_t(batch) * H * 8;
}

// Z = FFT(x + iy) = X + iY with X, Y the R2C spectra of the real fields x, y
// (extended to the full grid by X(-k) = conj X(k)); the inverse transform is
// conj(FFT(conj Z)).  Only R2C plans are used.
int fft_c2c(hipStream_t s, int nx, int ny, int nz, int batch, const void *in, void *out, int direction, void *ws,
            size_t ws_bytes) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(batch >= 1 && batch <= 32767, "batch %d out of range", batch);
  JPM_REQUIRE(direction == -1 || direction == 1, "direction must be -1 (forward) or +1 (inverse)");
  JPM_REQUIRE(in && out, "null FFT buffer");
  const size_t need = fft_c2c_workspace_size(nx, ny, nz, batch);
  JPM_REQUIRE(ws && ws_bytes >= need, "c2c FFT needs %zu workspace bytes, got %zu", need, ws_bytes);
  const int64_t M = cells(nx, ny, nz);
  float *xy = static_cast<float *>(ws);
  void *xyh = static_cast<char *>(ws) + align_up(2 * size_t(batch) * size_t(M) * 4);
  const float sign = direction == -1 ? 1.0f : -1.0f;
  rc = c2c_split(s, M, batch, in, xy, sign);
  if (rc) return rc;
  rc = fft_r2c(s, nx, ny, nz, 2 * batch, xy, xyh);
  if (rc) return rc;
  return c2c_combine(s, nx, ny, nz, batch, xyh, out, sign);
}

void free_twiddles();

void free_plans() {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto &kv : g_plans) hipfftDestroy(kv.second);
    g_plans.clear();
  }
  free_twiddles();
}

}  // namespace jpm
