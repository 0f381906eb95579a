// extern "C" entry points of libjaxpm_hip.so (declared in jaxpm_hip.h).
// Each wrapper clears the thread's error, forwards to the jpm:: routine and
// turns any C++ exception into JPM_E_NOMEM / JPM_E_INVALID.
#include <atomic>
#include <cstring>
#include <exception>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "tiles.hpp"

namespace jpm {

static thread_local std::string t_err;

int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  t_err = buf;
  return code;
}

void clear_error() { t_err.clear(); }

// ------------------------------------------------------------- timing
static std::atomic<bool> g_timing{false};
static std::mutex g_timing_mu;
struct TimedLaunch {
  const char *name;
  hipEvent_t a, b;
};
static std::vector<TimedLaunch> g_launches;

// Calls that must not allocate switch the timing off on their thread (its
// events are created per launch).
static thread_local int t_timing_paused = 0;
bool timing_enabled() { return t_timing_paused == 0 && g_timing.load(std::memory_order_relaxed); }
TimingPause::TimingPause() { ++t_timing_paused; }
TimingPause::~TimingPause() { --t_timing_paused; }

void timing_record(hipStream_t, const char *name, hipEvent_t begin, hipEvent_t end) {
  std::lock_guard<std::mutex> lk(g_timing_mu);
  g_launches.push_back({name, begin, end});
}
const char *last_error() { return t_err.c_str(); }

int cic_paint(hipStream_t, int, int, int, const float *, int64_t, const float *, float, float *, uint32_t,
              void *, size_t);
int cic_read(hipStream_t, int, int, int, const float *, int, int64_t, const float *, int64_t, float *,
             uint32_t);
int fft_r2c(hipStream_t, int, int, int, int, const float *, void *);
int fft_c2r(hipStream_t, int, int, int, int, void *, float *);
int fft_c2c(hipStream_t, int, int, int, int, const void *, void *, int, void *, size_t);
size_t fft_c2c_workspace_size(int, int, int, int);
void free_plans();
void free_tabs();
int noalloc_init_current_device();
int binned_paint_noalloc(hipStream_t, int, int, int, const float *, int64_t, const float *, float, int, float *,
                         uint32_t);
int kspace_force(hipStream_t, int, int, int, const void *, void *, float, float, uint32_t);
int kspace_compensate(hipStream_t, int, int, int, void *, float);
int kspace_linear_field(hipStream_t, int, int, int, void *, float, float, float, const float *, const float *,
                        int, float);
int kspace_shear6(hipStream_t, int, int, int, const void *, void *, float, uint32_t);
int lpt2_source(hipStream_t, int64_t, const float *, float *);
int cic_paint_f64(hipStream_t, int, int, int, const double *, int64_t, const double *, double, double *, uint32_t);
int cic_read_f64(hipStream_t, int, int, int, const double *, int, int64_t, const double *, int64_t, double *,
                 uint32_t);
size_t pm_forces_f64_workspace_size(int, int, int);
int pm_forces_f64(hipStream_t, int, int, int, const double *, int64_t, const double *, double, uint32_t, uint32_t,
                  double *, void *, size_t);
int cic_paint_multi(hipStream_t, int, int, int, const float *, int64_t, const float *, int, float *, uint32_t, void *,
                    size_t);
int cic_grad_pos(hipStream_t, int, int, int, const float *, int, int64_t, const float *, int64_t, const float *,
                 const float *, float, float *, uint32_t, int);
int kspace_force_adjoint(hipStream_t, int, int, int, const void *, void *, float, float, uint32_t);
int spectrum_full_to_half(hipStream_t, int, int, int, const void *, void *);
int kspace_shear6_adjoint(hipStream_t, int, int, int, const void *, void *, float, uint32_t);
int lpt2_source_vjp(hipStream_t, int64_t, const float *, const float *, float *);
int spectrum_half_to_full(hipStream_t, int, int, int, const void *, void *);
size_t power_spectrum_ws(int, int, int, int, int, int);
size_t linear_field_table_vjp_ws(int);
int kspace_linear_field_table_vjp(hipStream_t, int, int, int, const void *, const void *, float, float, float,
                                  const float *, const float *, int, float, double *, void *, size_t);
size_t dot_ws();
int batch_create(int, const int *, const int64_t *, uint32_t, SimBatch **);
int batch_destroy(SimBatch *);
size_t batch_kick_ws(const SimBatch &);
size_t batch_vjp_ws(const SimBatch &);
int batch_force_kick(hipStream_t, const SimBatch &, const float *, float *, const float *, float *, float, uint32_t,
                     int, void *, size_t, void *, size_t, const float *, float *, float *, const uint32_t *,
                     uint32_t *, float *);
int batch_force_vjp(hipStream_t, const SimBatch &, const float *, const float *, const float *, const float *,
                    const float *, float *, float, uint32_t, void *, size_t, const void *, size_t, const float *,
                    float *, const float *, float *);
int batch_axpy(hipStream_t, const SimBatch &, const float *, const float *, float *);
size_t batch_dot_ws(const SimBatch &);
int batch_dot(hipStream_t, const SimBatch &, const float *, const float *, double *, void *, size_t);
size_t batch_lpt_ws(const SimBatch &);
int batch_linear_field(hipStream_t, const SimBatch &, const float *, float *, const float *, const float *,
                       const float *, int, void *, size_t);
int batch_lattice_forces(hipStream_t, const SimBatch &, const float *, float *, void *, size_t);
int batch_lattice_forces_vjp(hipStream_t, const SimBatch &, const float *, float *, void *, size_t);
int batch_paint_mesh(hipStream_t, const SimBatch &, const float *, float *, void *, size_t);
int batch_tile_order(hipStream_t, const SimBatch &, const float *, uint32_t *, void *, size_t);
int batch_paint_vjp(hipStream_t, const SimBatch &, const float *, const float *, float *);
int dot(hipStream_t, int64_t, const float *, const float *, double *, void *, size_t);
int power_spectrum(hipStream_t, int, int, int, const void *, const void *, double, double, double, const double *,
                   int, const int *, int, double, double, double, double *, void *, size_t);

int axpy(hipStream_t, int64_t, float, const float *, float *);
int permute_rows(hipStream_t, int64_t, const uint32_t *, const float *, float *, const float *, float *, int);
int paint_2d(hipStream_t, int, int, const float *, int64_t, int, const float *, float *);
int paint_2d_vjp(hipStream_t, int, int, const float *, int64_t, int, const float *, const float *, float *, int,
                 float *);
int density_plane(hipStream_t, const float *, int64_t, int, int, double, double, int, float *);
int density_plane_vjp(hipStream_t, const float *, int64_t, int, int, double, double, int, float *, float *);
size_t gaussian_smoothing_ws(int, int);
int gaussian_smoothing(hipStream_t, int, int, const float *, float *, double, void *, size_t);
int kspace_potential(hipStream_t, int, int, int, const void *, void *, float, float, uint32_t);
int fd4_gradient(hipStream_t, int, int, int, const float *, float, float *);
int fd4_divergence(hipStream_t, int, int, int, const float *, float, float *);

// Force meshes phi3 = scale * -g_d G Lr [C] delta_k, inverse-transformed.
// With the default 4-point gradient this is one inverse FFT of the potential
// followed by the real-space stencil (`tmp_mesh`, M floats, holds the
// potential); with JPM_K_GRAD_EXACT it is three inverse FFTs of the three
// gradient spectra (`spec3`, 3 half spectra).
static int force_meshes(hipStream_t s, int nx, int ny, int nz, const void *delta_k, float r_split, float scale,
                        uint32_t kflags, void *spec3, float *tmp_mesh, float *phi3) {
  const int64_t m = cells(nx, ny, nz);
  const float sc = float(double(scale) / double(m));
  int rc;
  if (kflags & JPM_K_GRAD_EXACT) {
    rc = kspace_force(s, nx, ny, nz, delta_k, spec3, r_split, sc, kflags);
    if (rc) return rc;
    return fft_c2r(s, nx, ny, nz, 3, spec3, phi3);
  }
  rc = kspace_potential(s, nx, ny, nz, delta_k, spec3, r_split, sc, kflags);
  if (rc) return rc;
  rc = fft_c2r(s, nx, ny, nz, 1, spec3, tmp_mesh);
  if (rc) return rc;
  return fd4_gradient(s, nx, ny, nz, tmp_mesh, 1.0f, phi3);
}
int binned_paint(hipStream_t, int, int, int, const float *, int64_t, const float *, float, int, float *, uint32_t,
                 void *, size_t);
size_t binned_paint_workspace_size(int, int, int, int64_t, int, uint32_t);
bool binned_paint_supported(int, int, int);

// Paint used inside the pipelines: the binned LDS paint unless JPM_ATOMIC or
// the mesh has more tiles than the binned paint's LDS histogram holds.
// Pipeline flags: paint flags in bits 0..7, k-space flags in bits 8..15,
// sort and gather tuning in bits 16..25 (which travel with the paint flags).
static uint32_t paint_flags(uint32_t flags) { return flags & 0x07ff00ffu; }
static uint32_t kspace_flags(uint32_t flags) { return (flags >> 8) & 0xffu; }

static bool use_atomic(int nx, int ny, int nz, uint32_t pflags) {
  return (pflags & JPM_ATOMIC) || !binned_paint_supported(nx, ny, nz);
}

static size_t pipeline_paint_ws(int nx, int ny, int nz, int64_t n, int nc, uint32_t pflags) {
  if (use_atomic(nx, ny, nz, pflags))
    return (pflags & JPM_DETERMINISTIC) ? align_up(size_t(nc) * size_t(cells(nx, ny, nz)) * 8) + 256 : 0;
  return binned_paint_workspace_size(nx, ny, nz, n, nc, pflags);
}

static int pipeline_paint(hipStream_t s, int nx, int ny, int nz, const float *pos, int64_t n, const float *w, int nc,
                          float *mesh, uint32_t pflags, void *ws, size_t ws_bytes) {
  const uint32_t f = (pflags & (JPM_RELATIVE | JPM_DETERMINISTIC)) | JPM_ZERO_MESH;
  if (use_atomic(nx, ny, nz, pflags)) {
    if (nc == 1) return cic_paint(s, nx, ny, nz, pos, n, w, 1.0f, mesh, f, ws, ws_bytes);
    return cic_paint_multi(s, nx, ny, nz, pos, n, w, nc, mesh, f, ws, ws_bytes);
  }
  return binned_paint(s, nx, ny, nz, pos, n, w, 1.0f, nc, mesh, f, ws, ws_bytes);
}
int kick(hipStream_t, int, int, int, const float *, const float *, int64_t, float, float *, uint32_t);

size_t bins_paint_workspace_size(int, int, int, int64_t, int, uint32_t);
int bins_kick(hipStream_t, const Bins &, const float *, float, float *, int, float, float *, float *, uint32_t *,
              float *);
int bins_grad(hipStream_t, const Bins &, const float *, const float *, const float *, float, float *, bool, float,
              float *, float *);
int bins_read(hipStream_t, const Bins &, const float *, int, uint32_t, float *);

// ------------------------------------------------------ sorted pipelines
//
// The leapfrog's two hot calls, built on one tile sort of the positions:
//   force_kick: sort, window paint, R2C, potential, C2R -> phi (kept for the
//     adjoint), then the window gather applies the 4-point stencil to phi and
//     kicks v += gamma F.
//   force_vjp_pot: sort, 3-channel window paint of the cotangent, stencil
//     divergence, R2C, potential, C2R -> rho_bar, then one window gather for
//     both the direct term (on phi) and the density term (on rho_bar).

struct SortedWs {
  size_t sort, paint, mesh, g3, spec, spec2, faces, total;
  bool apart;   // the paint's face cells kept apart for the potential solve
};

// Absolute positions on a mesh of full tiles that the library's FFT takes:
// the paint hands its face cells to the potential solve's first pass in a
// buffer of their own (FaceCells) instead of writing them into the mesh,
// unless the flags ask for the mesh (JPM_TUNE_FACES_IN_MESH).
static bool faces_apart(int nx, int ny, int nz, uint32_t pflags) {
  return !(pflags & (JPM_RELATIVE | kTuneFacesInMesh)) && potential_faces_ok(nx, ny, nz);
}

static SortedWs sorted_ws(int nx, int ny, int nz, int64_t n, int nc, uint32_t pflags) {
  const size_t m = size_t(cells(nx, ny, nz)), h = size_t(half_cells(nx, ny, nz));
  SortedWs w;
  w.apart = nc == 1 && faces_apart(nx, ny, nz, pflags);
  const size_t tiles = w.apart ? size_t(nx / kTile) * size_t(ny / kTile) * size_t(nz / kTile) : 0;
  size_t o = 0;
  w.sort = o;  o += align_up(bin_sort_workspace_size(nx, ny, nz, n, pflags));
  w.paint = o; o += align_up(bins_paint_workspace_size(nx, ny, nz, n, nc, pflags));
  w.mesh = o;  o += align_up(m * 4);
  w.g3 = o;    o += (nc == 3) ? align_up(3 * m * 4) : 0;
  w.spec = o;  o += align_up(h * 8);
  w.spec2 = o; o += align_up(h * 8);
  w.faces = o; o += (nc == 1) ? align_up(tiles * kFaceStride * 4) : 0;
  w.total = o;
  return w;
}

static FaceCells sorted_faces(const SortedWs &w, char *base, int nx, int ny, int nz) {
  const size_t tiles = size_t(nx / kTile) * size_t(ny / kTile) * size_t(nz / kTile);
  return FaceCells{reinterpret_cast<float *>(base + w.faces), uint32_t(ny), uint32_t(ny / kTile),
                   uint32_t(nz / kTile), tiles * kFaceStride, FastDiv(uint32_t(ny))};
}

static bool sorted_supported(int nx, int ny, int nz, uint32_t flags) {
  return binned_paint_supported(nx, ny, nz) && !(kspace_flags(flags) & JPM_K_GRAD_EXACT) && !(flags & JPM_ATOMIC);
}

static int force_kick(hipStream_t s, int nx, int ny, int nz, const float *pos, int64_t n, float *vel, float gamma,
                      float *phi, float r_split, uint32_t flags, int write_forces, void *ws, size_t ws_bytes,
                      void *keep_sort, size_t keep_bytes, float beta = 0.0f, float *xout = nullptr,
                      float *xkeep = nullptr, const uint32_t *counts_in = nullptr, uint32_t *counts_out = nullptr,
                      float *vout = nullptr) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(!xout || !write_forces, "the fused drift follows a kick, not a force write");
  JPM_REQUIRE(!counts_out || (xout && !(flags & JPM_RELATIVE)), "tile counts of drifted absolute positions only");
  const uint32_t pflags = paint_flags(flags), kflags = kspace_flags(flags);
  JPM_REQUIRE(sorted_supported(nx, ny, nz, flags), "the sorted pipeline needs the 4-point gradient, the binned paint and 3 meshes within 2^32 cells");
  JPM_REQUIRE(n >= 1, "no particles");
  JPM_REQUIRE(pos && phi, "null pointer argument");
  JPM_REQUIRE(vel || !(write_forces || xout), "the force write and the fused drift need vel");
  const SortedWs w = sorted_ws(nx, ny, nz, n, 1, pflags);
  JPM_REQUIRE(ws && ws_bytes >= w.total, "force_kick needs %zu workspace bytes, got %zu", w.total, ws_bytes);
  char *base = static_cast<char *>(ws);
  const int64_t m = cells(nx, ny, nz);
  Bins b;
  if (keep_sort) {
    JPM_REQUIRE(keep_bytes >= bin_sort_workspace_size(nx, ny, nz, n, pflags), "keep_sort buffer too small (%zu bytes)",
                keep_bytes);
    rc = bin_sort(s, nx, ny, nz, pos, n, pflags, keep_sort, keep_bytes, b, counts_in);
  } else {
    rc = bin_sort(s, nx, ny, nz, pos, n, pflags, base + w.sort, w.paint - w.sort, b, counts_in);
  }
  if (rc) return rc;
  float *mesh = reinterpret_cast<float *>(base + w.mesh);
  const FaceCells fc = sorted_faces(w, base, nx, ny, nz);
  const FaceCells *fcp = w.apart ? &fc : nullptr;
  rc = bins_paint(s, b, nullptr, 1.0f, 1, mesh, JPM_ZERO_MESH, base + w.paint, w.mesh - w.paint, nullptr, fcp);
  if (rc) return rc;
  rc = potential_solve(s, nx, ny, nz, 1, mesh, base + w.spec, base + w.spec2, phi, r_split, float(1.0 / double(m)),
                       nullptr, kflags, fcp);
  if (rc) return rc;
  // without vel the call is the force evaluation alone: potential and sort
  if (!vel) return JPM_OK;
  return bins_kick(s, b, phi, gamma, vel, write_forces, beta, xout, xkeep, counts_out, vout);
}

// cic_read of up to 3 meshes through a tile sort of the positions (absolute
// mode): for particles in no spatial order, whose direct 8-corner gathers
// would each touch separate cache lines.
static int cic_read_sorted(hipStream_t s, int nx, int ny, int nz, const float *mesh, int ncomp, int64_t comp_stride,
                           const float *pos, int64_t n, float *out, uint32_t flags, void *ws, size_t ws_bytes) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(!(flags & JPM_RELATIVE), "the tile-sorted read takes absolute positions");
  JPM_REQUIRE(binned_paint_supported(nx, ny, nz), "mesh too large for the tile sort");
  JPM_REQUIRE(ncomp >= 1 && ncomp <= 3, "ncomp=%d must be 1..3", ncomp);
  JPM_REQUIRE(n >= 0 && n <= kMaxCells, "particle count %lld out of range", (long long)n);
  const int64_t m = cells(nx, ny, nz);
  JPM_REQUIRE(ncomp == 1 || (comp_stride >= m && int64_t(ncomp - 1) * comp_stride + m <= (int64_t(1) << 32)),
              "bad component stride %lld", (long long)comp_stride);
  if (n == 0) return JPM_OK;
  JPM_REQUIRE(mesh && pos && out, "null pointer argument");
  const size_t need = bin_sort_workspace_size(nx, ny, nz, n, 0);
  JPM_REQUIRE(ws && ws_bytes >= need, "tile-sorted read needs %zu workspace bytes, got %zu", need, ws_bytes);
  Bins b;
  rc = bin_sort(s, nx, ny, nz, pos, n, 0, ws, ws_bytes, b);
  if (rc) return rc;
  return bins_read(s, b, mesh, ncomp, uint32_t(ncomp == 1 ? m : comp_stride), out);
}

static int tile_order(hipStream_t s, int nx, int ny, int nz, const float *pos, int64_t n, uint32_t *order, void *ws,
                      size_t ws_bytes) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(binned_paint_supported(nx, ny, nz), "mesh too large for the tile sort");
  JPM_REQUIRE(n >= 0 && n <= kMaxCells, "particle count %lld out of range", (long long)n);
  if (n == 0) return JPM_OK;
  JPM_REQUIRE(pos && order, "null pointer argument");
  const size_t need = bin_sort_workspace_size(nx, ny, nz, n, 0);
  JPM_REQUIRE(ws && ws_bytes >= need, "tile order needs %zu workspace bytes, got %zu", need, ws_bytes);
  Bins b;
  rc = bin_sort(s, nx, ny, nz, pos, n, 0, ws, ws_bytes, b);
  if (rc) return rc;
  return bins_order(s, b, order);
}

static int force_vjp_pot(hipStream_t s, int nx, int ny, int nz, const float *phi, const float *pos, int64_t n,
                         const float *cot, float scale, float *gpos, float r_split, uint32_t flags, void *ws,
                         size_t ws_bytes, const void *sorted, size_t sorted_bytes, float beta = 0.0f,
                         float *cot_out = nullptr, const float *cot_max_in = nullptr, float *cot_max_out = nullptr) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  const uint32_t pflags = paint_flags(flags), kflags = kspace_flags(flags);
  JPM_REQUIRE(sorted_supported(nx, ny, nz, flags), "the sorted pipeline needs the 4-point gradient, the binned paint and 3 meshes within 2^32 cells");
  JPM_REQUIRE(n >= 1, "no particles");
  JPM_REQUIRE(phi && pos && cot && gpos, "null pointer argument");
  const SortedWs w = sorted_ws(nx, ny, nz, n, 3, pflags);
  JPM_REQUIRE(ws && ws_bytes >= w.total, "force_vjp needs %zu workspace bytes, got %zu", w.total, ws_bytes);
  char *base = static_cast<char *>(ws);
  const int64_t m = cells(nx, ny, nz);
  Bins b;
  if (sorted) {
    // the sort of these positions kept by jpm_force_kick_f32
    JPM_REQUIRE(sorted_bytes >= bin_sort_workspace_size(nx, ny, nz, n, pflags), "sorted buffer too small (%zu bytes)",
                sorted_bytes);
    b = bins_view(nx, ny, nz, n, pflags, sorted);
  } else {
    rc = bin_sort(s, nx, ny, nz, pos, n, pflags, base + w.sort, w.paint - w.sort, b);
    if (rc) return rc;
  }
  float *g3 = reinterpret_cast<float *>(base + w.g3);
  float *rho = reinterpret_cast<float *>(base + w.mesh);
  rc = bins_paint(s, b, cot, 1.0f, 3, g3, JPM_ZERO_MESH, base + w.paint, w.mesh - w.paint, cot_max_in);
  if (rc) return rc;
  rc = potential_solve_div(s, nx, ny, nz, 1, g3, size_t(m), rho, base + w.spec, base + w.spec2, rho, r_split,
                           float(double(scale) / double(m)), nullptr, kflags);
  if (rc) return rc;
  return bins_grad(s, b, phi, rho, cot, scale, gpos, tuned_rho_lds(flags), beta, cot_out, cot_max_out);
}

// Workspace of force_field: [mesh][half spectrum][3 half spectra][paint scratch].
struct FieldWs {
  size_t mesh, spec, spec3, paint, total;
};

static FieldWs field_ws(int nx, int ny, int nz, int64_t n, uint32_t flags) {
  const size_t m = size_t(cells(nx, ny, nz)), h = size_t(half_cells(nx, ny, nz));
  FieldWs w;
  w.mesh = 0;
  w.spec = align_up(m * sizeof(float));
  w.spec3 = w.spec + align_up(h * 8);
  w.paint = w.spec3 + align_up(3 * h * 8);
  w.total = w.paint + pipeline_paint_ws(nx, ny, nz, n, 1, flags);
  return w;
}

// phi3 = scale * Re IFFT[-g_d G Lr FFT(paint(pos) or delta_k)], the three
// force meshes that the read / kick sample.
static int force_field(hipStream_t s, int nx, int ny, int nz, const float *pos, int64_t n, const void *delta_k,
                       float *phi3, float r_split, float scale, uint32_t flags, void *ws, size_t ws_bytes) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  const uint32_t pflags = paint_flags(flags), kflags = kspace_flags(flags);
  const FieldWs w = field_ws(nx, ny, nz, n, pflags);
  JPM_REQUIRE(ws && ws_bytes >= w.total, "force_field needs %zu workspace bytes, got %zu", w.total, ws_bytes);
  JPM_REQUIRE(phi3 != nullptr, "phi3 is null");
  char *base = static_cast<char *>(ws);
  float *mesh = reinterpret_cast<float *>(base + w.mesh);
  void *spec = base + w.spec;
  void *spec3 = base + w.spec3;
  const int64_t m = cells(nx, ny, nz);
  if (!delta_k) {
    rc = pipeline_paint(s, nx, ny, nz, pos, n, nullptr, 1, mesh, pflags, base + w.paint, w.total - w.paint);
    if (rc) return rc;
    rc = fft_r2c(s, nx, ny, nz, 1, mesh, spec);
    if (rc) return rc;
    delta_k = spec;
  }
  return force_meshes(s, nx, ny, nz, delta_k, r_split, scale, kflags, spec3, mesh, phi3);
}

// Workspace of force_vjp: [3 cotangent meshes][3 half spectra][half
// spectrum][density cotangent mesh][paint scratch].
struct VjpWs {
  size_t g3, spec3, spec, rho, paint, total;
};

static VjpWs vjp_ws(int nx, int ny, int nz, int64_t n, uint32_t flags) {
  const size_t m = size_t(cells(nx, ny, nz)), h = size_t(half_cells(nx, ny, nz));
  VjpWs w;
  w.g3 = 0;
  w.spec3 = align_up(3 * m * sizeof(float));
  w.spec = w.spec3 + align_up(3 * h * 8);
  w.rho = w.spec + align_up(h * 8);
  w.paint = w.rho + align_up(m * sizeof(float));
  w.total = w.paint + pipeline_paint_ws(nx, ny, nz, n, 3, flags);
  return w;
}

// gpos += scale * J^T cot, J the Jacobian of the forces read at pos from the
// field of the particles at pos (pm.py:12-58): the direct term samples the
// gradient of phi3, the density term paints cot (the transpose of the read),
// applies the transposed k-space operator and samples the gradient of the
// resulting density cotangent.
static int force_vjp(hipStream_t s, int nx, int ny, int nz, const float *phi3, const float *pos, int64_t n,
                     const float *cot, float scale, float *gpos, float r_split, uint32_t flags, void *ws,
                     size_t ws_bytes) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  const uint32_t pflags = paint_flags(flags), kflags = kspace_flags(flags);
  const VjpWs w = vjp_ws(nx, ny, nz, n, pflags);
  JPM_REQUIRE(ws && ws_bytes >= w.total, "force_vjp needs %zu workspace bytes, got %zu", w.total, ws_bytes);
  JPM_REQUIRE(phi3 && pos && cot && gpos, "null pointer argument");
  char *base = static_cast<char *>(ws);
  float *g3 = reinterpret_cast<float *>(base + w.g3);
  void *spec3 = base + w.spec3;
  void *spec = base + w.spec;
  float *rho = reinterpret_cast<float *>(base + w.rho);
  const int64_t m = cells(nx, ny, nz);
  const uint32_t rel = pflags & JPM_RELATIVE;
  rc = pipeline_paint(s, nx, ny, nz, pos, n, cot, 3, g3, pflags, base + w.paint, w.total - w.paint);
  if (rc) return rc;
  if (kflags & JPM_K_GRAD_EXACT) {
    rc = fft_r2c(s, nx, ny, nz, 3, g3, spec3);
    if (rc) return rc;
    rc = kspace_force_adjoint(s, nx, ny, nz, spec3, spec, r_split, float(double(scale) / double(m)), kflags);
    if (rc) return rc;
  } else {
    // rho_bar = A (sum_d D_d g_d): the stencil's transpose in real space,
    // then the (self-adjoint) potential operator A.
    rc = fd4_divergence(s, nx, ny, nz, g3, 1.0f, rho);
    if (rc) return rc;
    rc = fft_r2c(s, nx, ny, nz, 1, rho, spec3);
    if (rc) return rc;
    rc = kspace_potential(s, nx, ny, nz, spec3, spec, r_split, float(double(scale) / double(m)), kflags);
    if (rc) return rc;
  }
  rc = fft_c2r(s, nx, ny, nz, 1, spec, rho);
  if (rc) return rc;
  rc = cic_grad_pos(s, nx, ny, nz, phi3, 3, m, pos, n, cot, nullptr, scale, gpos, rel, 1);
  if (rc) return rc;
  return cic_grad_pos(s, nx, ny, nz, rho, 1, m, pos, n, nullptr, nullptr, 1.0f, gpos, rel, 1);
}

// Workspace layout of pm_forces: [3 meshes][1 half spectrum][3 half
// spectra][deterministic-paint scratch].  The painted mesh lives in the first
// of the three force meshes; it is dead once the R2C has read it.
struct ForceWs {
  size_t meshes, spec, spec3, tmp, paint, total;
};

static ForceWs force_ws(int nx, int ny, int nz, int64_t n, uint32_t flags) {
  const size_t m = size_t(cells(nx, ny, nz)), h = size_t(half_cells(nx, ny, nz));
  ForceWs w;
  w.meshes = 0;
  w.spec = w.meshes + align_up(3 * m * sizeof(float));
  w.spec3 = w.spec + align_up(h * 8);
  w.tmp = w.spec3 + align_up(3 * h * 8);
  w.paint = w.tmp + align_up(m * sizeof(float));
  w.total = w.paint + pipeline_paint_ws(nx, ny, nz, n, 1, flags);
  return w;
}

static int pm_forces(hipStream_t s, int nx, int ny, int nz, const float *pos, int64_t n, const void *delta_k,
                     float *forces, float r_split, float scale, uint32_t flags, void *ws, size_t ws_bytes) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  const uint32_t pflags = paint_flags(flags), kflags = kspace_flags(flags);
  const ForceWs w = force_ws(nx, ny, nz, n, pflags);
  JPM_REQUIRE(ws && ws_bytes >= w.total, "pm_forces needs %zu workspace bytes, got %zu", w.total, ws_bytes);
  JPM_REQUIRE(forces != nullptr, "forces is null");
  char *base = static_cast<char *>(ws);
  float *meshes = reinterpret_cast<float *>(base + w.meshes);
  void *spec = base + w.spec;
  void *spec3 = base + w.spec3;
  const int64_t m = cells(nx, ny, nz);
  if (!delta_k) {
    rc = pipeline_paint(s, nx, ny, nz, pos, n, nullptr, 1, meshes, pflags, base + w.paint, w.total - w.paint);
    if (rc) return rc;
    rc = fft_r2c(s, nx, ny, nz, 1, meshes, spec);
    if (rc) return rc;
    delta_k = spec;
  }
  rc = force_meshes(s, nx, ny, nz, delta_k, r_split, scale, kflags, spec3, reinterpret_cast<float *>(base + w.tmp),
                    meshes);
  if (rc) return rc;
  return cic_read(s, nx, ny, nz, meshes, 3, m, pos, n, forces, pflags & JPM_RELATIVE);
}

}  // namespace jpm

using namespace jpm;

#define JPM_GUARD(expr)                                                         \
  do {                                                                          \
    clear_error();                                                              \
    try {                                                                       \
      return (expr);                                                            \
    } catch (const std::bad_alloc &) {                                          \
      return fail(JPM_E_NOMEM, "host allocation failed");                       \
    } catch (const std::exception &ex) {                                        \
      return fail(JPM_E_INVALID, "internal error: %s", ex.what());              \
    } catch (...) {                                                             \
      return fail(JPM_E_INVALID, "internal error");                             \
    }                                                                           \
  } while (0)

extern "C" {

int jpm_init(void) {
  clear_error();
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(JPM_E_HIP, "no HIP device available");
  // the allocation-free paint's scratch address on the current device
  return noalloc_init_current_device();
}

int jpm_shutdown(void) {
  clear_error();
  free_plans();
  free_tabs();
  return JPM_OK;
}

int jpm_abi_version(void) { return JPM_ABI_VERSION; }

int jpm_timing_enable(int on) {
  clear_error();
  g_timing.store(on != 0);
  return JPM_OK;
}

int jpm_timing_report(char *buf, size_t len) {
  clear_error();
  std::vector<TimedLaunch> ls;
  {
    std::lock_guard<std::mutex> lk(g_timing_mu);
    ls.swap(g_launches);
  }
  std::map<std::string, std::pair<long, double>> agg;
  int rc = JPM_OK;
  for (auto &l : ls) {
    float ms = 0.0f;
    if (hipEventSynchronize(l.b) == hipSuccess && hipEventElapsedTime(&ms, l.a, l.b) == hipSuccess) {
      auto &e = agg[l.name];
      e.first += 1;
      e.second += ms;
    } else {
      rc = fail(JPM_E_HIP, "timing event failed for %s", l.name);
    }
    (void)hipEventDestroy(l.a);
    (void)hipEventDestroy(l.b);
  }
  std::string out;
  for (auto &kv : agg) {
    char line[256];
    snprintf(line, sizeof line, "%s %ld %.6f\n", kv.first.c_str(), kv.second.first, kv.second.second);
    out += line;
  }
  if (buf && len) {
    strncpy(buf, out.c_str(), len - 1);
    buf[len - 1] = 0;
    if (out.size() >= len) return rc ? rc : fail(JPM_E_NOMEM, "timing report truncated (%zu bytes)", out.size());
  }
  return rc;
}

const char *jpm_last_error(void) { return last_error(); }

size_t jpm_cic_paint_workspace_size(int nx, int ny, int nz, int64_t n, uint32_t flags) {
  (void)n;
  if (!(flags & JPM_DETERMINISTIC) || nx <= 0 || ny <= 0 || nz <= 0) return 0;
  return align_up(size_t(cells(nx, ny, nz)) * 8) + 256;
}

int jpm_cic_paint_f32(void *stream, int nx, int ny, int nz, const float *pos, int64_t n, const float *weight,
                      float wscalar, float *mesh, uint32_t flags, void *ws, size_t ws_bytes) {
  JPM_GUARD(cic_paint(as_stream(stream), nx, ny, nz, pos, n, weight, wscalar, mesh, flags, ws, ws_bytes));
}

int jpm_cic_read_f32(void *stream, int nx, int ny, int nz, const float *mesh, int ncomp, int64_t comp_stride,
                     const float *pos, int64_t n, float *out, uint32_t flags) {
  JPM_GUARD(cic_read(as_stream(stream), nx, ny, nz, mesh, ncomp, comp_stride, pos, n, out, flags));
}

int jpm_cic_paint_f64(void *stream, int nx, int ny, int nz, const double *pos, int64_t n, const double *weight,
                      double wscalar, double *mesh, uint32_t flags) {
  JPM_GUARD(cic_paint_f64(as_stream(stream), nx, ny, nz, pos, n, weight, wscalar, mesh, flags));
}

int jpm_cic_read_f64(void *stream, int nx, int ny, int nz, const double *mesh, int ncomp, int64_t comp_stride,
                     const double *pos, int64_t n, double *out, uint32_t flags) {
  JPM_GUARD(cic_read_f64(as_stream(stream), nx, ny, nz, mesh, ncomp, comp_stride, pos, n, out, flags));
}

size_t jpm_pm_forces_f64_workspace_size(int nx, int ny, int nz) { return pm_forces_f64_workspace_size(nx, ny, nz); }

int jpm_pm_forces_f64(void *stream, int nx, int ny, int nz, const double *pos, int64_t n, const double *delta,
                      double r_split, uint32_t flags, uint32_t kflags, double *out3, void *ws, size_t ws_bytes) {
  JPM_GUARD(pm_forces_f64(as_stream(stream), nx, ny, nz, pos, n, delta, r_split, flags, kflags, out3, ws, ws_bytes));
}

int jpm_fft_r2c_f32(void *stream, int nx, int ny, int nz, int batch, const float *in, void *out) {
  JPM_GUARD(fft_r2c(as_stream(stream), nx, ny, nz, batch, in, out));
}

int jpm_fft_c2r_f32(void *stream, int nx, int ny, int nz, int batch, void *in, float *out) {
  JPM_GUARD(fft_c2r(as_stream(stream), nx, ny, nz, batch, in, out));
}

size_t jpm_fft_c2c_workspace_size(int nx, int ny, int nz, int batch) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || batch < 1) return 0;
  return fft_c2c_workspace_size(nx, ny, nz, batch);
}

int jpm_fft_c2c_f32(void *stream, int nx, int ny, int nz, int batch, const void *in, void *out, int direction,
                    void *ws, size_t ws_bytes) {
  JPM_GUARD(fft_c2c(as_stream(stream), nx, ny, nz, batch, in, out, direction, ws, ws_bytes));
}

int jpm_kspace_force_f32(void *stream, int nx, int ny, int nz, const void *delta_k, void *out3, float r_split,
                         float scale, uint32_t flags) {
  JPM_GUARD(kspace_force(as_stream(stream), nx, ny, nz, delta_k, out3, r_split, scale, flags));
}

int jpm_kspace_potential_f32(void *stream, int nx, int ny, int nz, const void *delta_k, void *out, float r_split,
                             float scale, uint32_t flags) {
  JPM_GUARD(kspace_potential(as_stream(stream), nx, ny, nz, delta_k, out, r_split, scale, flags));
}

int jpm_fd4_gradient_f32(void *stream, int nx, int ny, int nz, const float *phi, float scale, float *out3) {
  JPM_GUARD(fd4_gradient(as_stream(stream), nx, ny, nz, phi, scale, out3));
}

int jpm_fd4_divergence_f32(void *stream, int nx, int ny, int nz, const float *g3, float scale, float *out) {
  JPM_GUARD(fd4_divergence(as_stream(stream), nx, ny, nz, g3, scale, out));
}

int jpm_kspace_compensate_f32(void *stream, int nx, int ny, int nz, void *field_k, float scale) {
  JPM_GUARD(kspace_compensate(as_stream(stream), nx, ny, nz, field_k, scale));
}

int jpm_kspace_linear_field_f32(void *stream, int nx, int ny, int nz, void *field_k, float Lx, float Ly, float Lz,
                                const float *k_tab, const float *pk_tab, int ntab, float scale) {
  JPM_GUARD(kspace_linear_field(as_stream(stream), nx, ny, nz, field_k, Lx, Ly, Lz, k_tab, pk_tab, ntab, scale));
}

int jpm_kspace_shear6_f32(void *stream, int nx, int ny, int nz, const void *delta_k, void *out6, float scale,
                          uint32_t flags) {
  JPM_GUARD(kspace_shear6(as_stream(stream), nx, ny, nz, delta_k, out6, scale, flags));
}

int jpm_lpt2_source_f32(void *stream, int64_t M, const float *phi6, float *delta2) {
  JPM_GUARD(lpt2_source(as_stream(stream), M, phi6, delta2));
}

size_t jpm_cic_paint_multi_workspace_size(int nx, int ny, int nz, int nc, uint32_t flags) {
  if (!(flags & JPM_DETERMINISTIC) || nx <= 0 || ny <= 0 || nz <= 0 || nc <= 0) return 0;
  return align_up(size_t(nc) * size_t(cells(nx, ny, nz)) * 8) + 256;
}

int jpm_cic_paint_multi_f32(void *stream, int nx, int ny, int nz, const float *pos, int64_t n,
                            const float *weights, int nc, float *mesh, uint32_t flags, void *ws, size_t ws_bytes) {
  JPM_GUARD(cic_paint_multi(as_stream(stream), nx, ny, nz, pos, n, weights, nc, mesh, flags, ws, ws_bytes));
}

int jpm_cic_grad_pos_f32(void *stream, int nx, int ny, int nz, const float *mesh, int ncomp, int64_t comp_stride,
                         const float *pos, int64_t n, const float *cot, const float *weight, float wscalar,
                         float *gpos, uint32_t flags, int accumulate) {
  JPM_GUARD(cic_grad_pos(as_stream(stream), nx, ny, nz, mesh, ncomp, comp_stride, pos, n, cot, weight, wscalar, gpos,
                         flags, accumulate));
}

int jpm_kspace_force_adjoint_f32(void *stream, int nx, int ny, int nz, const void *in3, void *out, float r_split,
                                 float scale, uint32_t flags) {
  JPM_GUARD(kspace_force_adjoint(as_stream(stream), nx, ny, nz, in3, out, r_split, scale, flags));
}

int jpm_kspace_shear6_adjoint_f32(void *stream, int nx, int ny, int nz, const void *in6, void *out, float scale,
                                  uint32_t flags) {
  JPM_GUARD(kspace_shear6_adjoint(as_stream(stream), nx, ny, nz, in6, out, scale, flags));
}

int jpm_lpt2_source_vjp_f32(void *stream, int64_t M, const float *phi6, const float *gbar, float *out6) {
  JPM_GUARD(lpt2_source_vjp(as_stream(stream), M, phi6, gbar, out6));
}

int jpm_spectrum_full_to_half_f32(void *stream, int nx, int ny, int nz, const void *full, void *half) {
  JPM_GUARD(spectrum_full_to_half(as_stream(stream), nx, ny, nz, full, half));
}

int jpm_spectrum_half_to_full_f32(void *stream, int nx, int ny, int nz, const void *half, void *full) {
  JPM_GUARD(spectrum_half_to_full(as_stream(stream), nx, ny, nz, half, full));
}

size_t jpm_linear_field_table_vjp_workspace_size(int ntab) { return linear_field_table_vjp_ws(ntab); }

int jpm_kspace_linear_field_table_vjp_f32(void *stream, int nx, int ny, int nz, const void *noise_k, const void *cot_k,
                                          float Lx, float Ly, float Lz, const float *k_tab, const float *pk_tab,
                                          int ntab, float scale, double *out, void *ws, size_t ws_bytes) {
  JPM_GUARD(kspace_linear_field_table_vjp(as_stream(stream), nx, ny, nz, noise_k, cot_k, Lx, Ly, Lz, k_tab, pk_tab,
                                          ntab, scale, out, ws, ws_bytes));
}

size_t jpm_dot_workspace_size(void) { return dot_ws(); }

int jpm_batch_create(int nsims, const int *dims, const int64_t *counts, uint32_t flags, void **batch) {
  JPM_GUARD(batch_create(nsims, dims, counts, flags, reinterpret_cast<SimBatch **>(batch)));
}

int jpm_batch_destroy(void *batch) { JPM_GUARD(batch_destroy(static_cast<SimBatch *>(batch))); }

static const SimBatch *as_batch(const void *b) { return static_cast<const SimBatch *>(b); }

size_t jpm_batch_sort_size(const void *batch) { return batch ? batch_sort_workspace_size(*as_batch(batch)) : 0; }
size_t jpm_batch_force_kick_workspace_size(const void *batch) { return batch ? batch_kick_ws(*as_batch(batch)) : 0; }
size_t jpm_batch_force_vjp_workspace_size(const void *batch) { return batch ? batch_vjp_ws(*as_batch(batch)) : 0; }
size_t jpm_batch_dot_workspace_size(const void *batch) { return batch ? batch_dot_ws(*as_batch(batch)) : 0; }

int jpm_batch_force_kick_f32(void *stream, const void *batch, const float *pos, float *vel, const float *gamma,
                             float *phi, float r_split, uint32_t kflags, int write_forces, void *ws, size_t ws_bytes,
                             void *keep_sort, size_t keep_bytes, const uint32_t *counts_in) {
  if (!batch) return fail(JPM_E_INVALID, "null batch");
  JPM_GUARD(batch_force_kick(as_stream(stream), *as_batch(batch), pos, vel, gamma, phi, r_split, kflags, write_forces,
                             ws, ws_bytes, keep_sort, keep_bytes, nullptr, nullptr, nullptr, counts_in, nullptr,
                             nullptr));
}

int jpm_batch_force_kick_drift_f32(void *stream, const void *batch, const float *pos, float *vel, const float *gamma,
                                   const float *beta, float *pos_out, float *pos_keep, float *phi, float r_split,
                                   uint32_t kflags, void *ws, size_t ws_bytes, void *keep_sort, size_t keep_bytes,
                                   const uint32_t *counts_in, uint32_t *counts_out, float *vel_out) {
  if (!batch) return fail(JPM_E_INVALID, "null batch");
  if (!pos_out || !beta) return fail(JPM_E_INVALID, "pos_out and beta are required");
  JPM_GUARD(batch_force_kick(as_stream(stream), *as_batch(batch), pos, vel, gamma, phi, r_split, kflags, 0, ws,
                             ws_bytes, keep_sort, keep_bytes, beta, pos_out, pos_keep, counts_in, counts_out,
                             vel_out));
}

size_t jpm_batch_tile_count(const void *batch) { return batch ? as_batch(batch)->K : 0; }

int jpm_batch_force_vjp_f32(void *stream, const void *batch, const float *phi, const float *pos, const float *cot,
                            const float *scale, const float *pot_scale, float *gpos, float r_split, uint32_t kflags,
                            void *ws, size_t ws_bytes, const void *sorted, size_t sorted_bytes) {
  if (!batch) return fail(JPM_E_INVALID, "null batch");
  JPM_GUARD(batch_force_vjp(as_stream(stream), *as_batch(batch), phi, pos, cot, scale, pot_scale, gpos, r_split,
                            kflags, ws, ws_bytes, sorted, sorted_bytes, nullptr, nullptr, nullptr, nullptr));
}

int jpm_batch_force_vjp_drift_f32(void *stream, const void *batch, const float *phi, const float *pos, float *cot,
                                  const float *scale, const float *pot_scale, const float *beta, float *gpos,
                                  float r_split, uint32_t kflags, void *ws, size_t ws_bytes, const void *sorted,
                                  size_t sorted_bytes, const float *cot_max_in, float *cot_max_out) {
  if (!batch) return fail(JPM_E_INVALID, "null batch");
  if (!beta || !cot) return fail(JPM_E_INVALID, "beta and cot are required");
  JPM_GUARD(batch_force_vjp(as_stream(stream), *as_batch(batch), phi, pos, cot, scale, pot_scale, gpos, r_split,
                            kflags, ws, ws_bytes, sorted, sorted_bytes, beta, cot, cot_max_in, cot_max_out));
}

int jpm_batch_axpy_f32(void *stream, const void *batch, const float *a, const float *x, float *y) {
  if (!batch) return fail(JPM_E_INVALID, "null batch");
  JPM_GUARD(batch_axpy(as_stream(stream), *as_batch(batch), a, x, y));
}

size_t jpm_batch_lpt_workspace_size(const void *batch) { return batch ? batch_lpt_ws(*as_batch(batch)) : 0; }

int jpm_batch_linear_field_f32(void *stream, const void *batch, const float *noise, float *out, const float *boxes,
                               const float *k_tab, const float *pk_tabs, int ntab, void *ws, size_t ws_bytes) {
  if (!batch) return fail(JPM_E_INVALID, "null batch");
  JPM_GUARD(batch_linear_field(as_stream(stream), *as_batch(batch), noise, out, boxes, k_tab, pk_tabs, ntab, ws,
                               ws_bytes));
}

int jpm_batch_lattice_forces_f32(void *stream, const void *batch, const float *field, float *psi, void *ws,
                                 size_t ws_bytes) {
  if (!batch) return fail(JPM_E_INVALID, "null batch");
  JPM_GUARD(batch_lattice_forces(as_stream(stream), *as_batch(batch), field, psi, ws, ws_bytes));
}

int jpm_batch_lattice_forces_vjp_f32(void *stream, const void *batch, const float *gpsi, float *gfield, void *ws,
                                     size_t ws_bytes) {
  if (!batch) return fail(JPM_E_INVALID, "null batch");
  JPM_GUARD(batch_lattice_forces_vjp(as_stream(stream), *as_batch(batch), gpsi, gfield, ws, ws_bytes));
}

int jpm_batch_tile_order_u32(void *stream, const void *batch, const float *pos, uint32_t *order, void *ws,
                             size_t ws_bytes) {
  JPM_REQUIRE(batch, "null batch handle");
  JPM_GUARD(batch_tile_order(as_stream(stream), *as_batch(batch), pos, order, ws, ws_bytes));
}

int jpm_batch_paint_f32(void *stream, const void *batch, const float *pos, float *mesh, void *ws, size_t ws_bytes) {
  if (!batch) return fail(JPM_E_INVALID, "null batch");
  JPM_GUARD(batch_paint_mesh(as_stream(stream), *as_batch(batch), pos, mesh, ws, ws_bytes));
}

int jpm_batch_paint_vjp_f32(void *stream, const void *batch, const float *pos, const float *gmesh, float *gpos) {
  if (!batch) return fail(JPM_E_INVALID, "null batch");
  JPM_GUARD(batch_paint_vjp(as_stream(stream), *as_batch(batch), pos, gmesh, gpos));
}

int jpm_batch_dot_f32(void *stream, const void *batch, const float *x, const float *y, double *out, void *ws,
                      size_t ws_bytes) {
  if (!batch) return fail(JPM_E_INVALID, "null batch");
  JPM_GUARD(batch_dot(as_stream(stream), *as_batch(batch), x, y, out, ws, ws_bytes));
}

int jpm_dot_f32(void *stream, int64_t n, const float *x, const float *y, double *out, void *ws, size_t ws_bytes) {
  JPM_GUARD(dot(as_stream(stream), n, x, y, out, ws, ws_bytes));
}

size_t jpm_power_spectrum_workspace_size(int nx, int ny, int nz, int nedges, int npoles, int cross) {
  return power_spectrum_ws(nx, ny, nz, nedges, npoles, cross);
}

int jpm_power_spectrum_f32(void *stream, int nx, int ny, int nz, const void *spec1, const void *spec2, double fac_x,
                           double fac_y, double fac_z, const double *edges, int nedges, const int *poles, int npoles,
                           double los_x, double los_y, double los_z, double *out, void *ws, size_t ws_bytes) {
  JPM_GUARD(power_spectrum(as_stream(stream), nx, ny, nz, spec1, spec2, fac_x, fac_y, fac_z, edges, nedges, poles,
                           npoles, los_x, los_y, los_z, out, ws, ws_bytes));
}

size_t jpm_force_field_workspace_size(int nx, int ny, int nz, int64_t n, uint32_t flags) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
  return field_ws(nx, ny, nz, n, paint_flags(flags)).total;
}

int jpm_force_field_f32(void *stream, int nx, int ny, int nz, const float *pos, int64_t n, const void *delta_k,
                        float *phi3, float r_split, float scale, uint32_t flags, void *ws, size_t ws_bytes) {
  JPM_GUARD(force_field(as_stream(stream), nx, ny, nz, pos, n, delta_k, phi3, r_split, scale, flags, ws, ws_bytes));
}

size_t jpm_force_vjp_workspace_size(int nx, int ny, int nz, int64_t n, uint32_t flags) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
  return vjp_ws(nx, ny, nz, n, paint_flags(flags)).total;
}

int jpm_force_vjp_f32(void *stream, int nx, int ny, int nz, const float *phi3, const float *pos, int64_t n,
                      const float *cot, float scale, float *gpos, float r_split, uint32_t flags, void *ws,
                      size_t ws_bytes) {
  JPM_GUARD(force_vjp(as_stream(stream), nx, ny, nz, phi3, pos, n, cot, scale, gpos, r_split, flags, ws, ws_bytes));
}

int jpm_axpy_f32(void *stream, int64_t n, float a, const float *x, float *y) {
  JPM_GUARD(axpy(as_stream(stream), n, a, x, y));
}

int jpm_cic_paint_2d_f32(void *stream, int nx, int ny, const float *pos, int64_t n, int stride,
                         const float *weight, float *mesh) {
  JPM_GUARD(paint_2d(as_stream(stream), nx, ny, pos, n, stride, weight, mesh));
}

int jpm_cic_paint_2d_vjp_f32(void *stream, int nx, int ny, const float *pos, int64_t n, int stride,
                             const float *weight, const float *g, float *gpos, int gstride, float *gweight) {
  JPM_GUARD(paint_2d_vjp(as_stream(stream), nx, ny, pos, n, stride, weight, g, gpos, gstride, gweight));
}

int jpm_density_plane_f32(void *stream, const float *pos, int64_t n, int box_nx, int box_ny, double center,
                          double width, int res, float *plane) {
  JPM_GUARD(density_plane(as_stream(stream), pos, n, box_nx, box_ny, center, width, res, plane));
}

int jpm_density_plane_vjp_f32(void *stream, const float *pos, int64_t n, int box_nx, int box_ny, double center,
                              double width, int res, float *g, float *gpos) {
  JPM_GUARD(density_plane_vjp(as_stream(stream), pos, n, box_nx, box_ny, center, width, res, g, gpos));
}

size_t jpm_gaussian_smoothing_workspace_size(int n0, int n1) { return gaussian_smoothing_ws(n0, n1); }

int jpm_gaussian_smoothing_f32(void *stream, int n0, int n1, const float *im, float *out, double sigma, void *ws,
                               size_t ws_bytes) {
  JPM_GUARD(gaussian_smoothing(as_stream(stream), n0, n1, im, out, sigma, ws, ws_bytes));
}

size_t jpm_tile_order_workspace_size(int nx, int ny, int nz, int64_t n) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !binned_paint_supported(nx, ny, nz)) return 0;
  return bin_sort_workspace_size(nx, ny, nz, n, 0);
}

int jpm_tile_order_u32(void *stream, int nx, int ny, int nz, const float *pos, int64_t n, uint32_t *order, void *ws,
                       size_t ws_bytes) {
  JPM_GUARD(tile_order(as_stream(stream), nx, ny, nz, pos, n, order, ws, ws_bytes));
}

int jpm_permute_rows_f32(void *stream, int64_t n, const uint32_t *order, const float *in0, float *out0,
                         const float *in1, float *out1, int scatter) {
  JPM_GUARD(permute_rows(as_stream(stream), n, order, in0, out0, in1, out1, scatter));
}

int jpm_kick_f32(void *stream, int nx, int ny, int nz, const float *phi3, const float *pos, int64_t n, float gamma,
                 float *vel, uint32_t flags) {
  JPM_GUARD(kick(as_stream(stream), nx, ny, nz, phi3, pos, n, gamma, vel, flags));
}

size_t jpm_pm_forces_workspace_size(int nx, int ny, int nz, int64_t n, uint32_t flags) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
  return force_ws(nx, ny, nz, n, paint_flags(flags)).total;
}

size_t jpm_force_kick_workspace_size(int nx, int ny, int nz, int64_t n, uint32_t flags) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !sorted_supported(nx, ny, nz, flags)) return 0;
  return sorted_ws(nx, ny, nz, n, 1, paint_flags(flags)).total;
}

int jpm_force_kick_f32(void *stream, int nx, int ny, int nz, const float *pos, int64_t n, float *vel, float gamma,
                       float *phi, float r_split, uint32_t flags, int write_forces, void *ws, size_t ws_bytes,
                       void *keep_sort, size_t keep_bytes, const uint32_t *counts_in) {
  JPM_GUARD(force_kick(as_stream(stream), nx, ny, nz, pos, n, vel, gamma, phi, r_split, flags, write_forces, ws,
                       ws_bytes, keep_sort, keep_bytes, 0.0f, nullptr, nullptr, counts_in, nullptr));
}

size_t jpm_bin_sort_size(int nx, int ny, int nz, int64_t n, uint32_t flags) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
  return bin_sort_workspace_size(nx, ny, nz, n, paint_flags(flags));
}

size_t jpm_force_vjp_pot_workspace_size(int nx, int ny, int nz, int64_t n, uint32_t flags) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !sorted_supported(nx, ny, nz, flags)) return 0;
  return sorted_ws(nx, ny, nz, n, 3, paint_flags(flags)).total;
}

int jpm_force_vjp_pot_f32(void *stream, int nx, int ny, int nz, const float *phi, const float *pos, int64_t n,
                          const float *cot, float scale, float *gpos, float r_split, uint32_t flags, void *ws,
                          size_t ws_bytes, const void *sorted, size_t sorted_bytes) {
  JPM_GUARD(force_vjp_pot(as_stream(stream), nx, ny, nz, phi, pos, n, cot, scale, gpos, r_split, flags, ws,
                          ws_bytes, sorted, sorted_bytes));
}

int jpm_force_kick_drift_f32(void *stream, int nx, int ny, int nz, const float *pos, int64_t n, float *vel,
                             float gamma, float beta, float *pos_out, float *pos_keep, float *phi, float r_split,
                             uint32_t flags, void *ws, size_t ws_bytes, void *keep_sort, size_t keep_bytes,
                             const uint32_t *counts_in, uint32_t *counts_out, float *vel_out) {
  if (!pos_out) return fail(JPM_E_INVALID, "pos_out is null");
  JPM_GUARD(force_kick(as_stream(stream), nx, ny, nz, pos, n, vel, gamma, phi, r_split, flags, 0, ws, ws_bytes,
                       keep_sort, keep_bytes, beta, pos_out, pos_keep, counts_in, counts_out, vel_out));
}

size_t jpm_tile_count(int nx, int ny, int nz) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
  return make_tiles(nx, ny, nz).K;
}

int jpm_force_vjp_pot_drift_f32(void *stream, int nx, int ny, int nz, const float *phi, const float *pos, int64_t n,
                                float *cot, float scale, float beta, float *gpos, float r_split, uint32_t flags,
                                void *ws, size_t ws_bytes, const void *sorted, size_t sorted_bytes,
                                const float *cot_max_in, float *cot_max_out) {
  if (!cot) return fail(JPM_E_INVALID, "cot is null");
  JPM_GUARD(force_vjp_pot(as_stream(stream), nx, ny, nz, phi, pos, n, cot, scale, gpos, r_split, flags, ws,
                          ws_bytes, sorted, sorted_bytes, beta, cot, cot_max_in, cot_max_out));
}

size_t jpm_cic_read_sorted_workspace_size(int nx, int ny, int nz, int64_t n) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !binned_paint_supported(nx, ny, nz)) return 0;
  return bin_sort_workspace_size(nx, ny, nz, n, 0);
}

int jpm_cic_read_sorted_f32(void *stream, int nx, int ny, int nz, const float *mesh, int ncomp, int64_t comp_stride,
                            const float *pos, int64_t n, float *out, uint32_t flags, void *ws, size_t ws_bytes) {
  JPM_GUARD(cic_read_sorted(as_stream(stream), nx, ny, nz, mesh, ncomp, comp_stride, pos, n, out, flags, ws,
                            ws_bytes));
}

size_t jpm_cic_paint_binned_workspace_size(int nx, int ny, int nz, int64_t n, int nc, uint32_t flags) {
  return binned_paint_workspace_size(nx, ny, nz, n, nc, flags);
}

int jpm_cic_paint_binned_f32(void *stream, int nx, int ny, int nz, const float *pos, int64_t n, const float *weights,
                             float wscalar, int nc, float *mesh, uint32_t flags, void *ws, size_t ws_bytes) {
  JPM_GUARD(binned_paint(as_stream(stream), nx, ny, nz, pos, n, weights, wscalar, nc, mesh, flags, ws, ws_bytes));
}

int jpm_pm_forces_f32(void *stream, int nx, int ny, int nz, const float *pos, int64_t n, const void *delta_k,
                      float *forces, float r_split, float scale, uint32_t flags, void *ws, size_t ws_bytes) {
  JPM_GUARD(pm_forces(as_stream(stream), nx, ny, nz, pos, n, delta_k, forces, r_split, scale, flags, ws, ws_bytes));
}

int jpm_cic_paint_binned_noalloc_f32(void *stream, int nx, int ny, int nz, const float *pos, int64_t n,
                                     const float *weights, float wscalar, int nc, float *mesh, uint32_t flags) {
  JPM_GUARD(binned_paint_noalloc(as_stream(stream), nx, ny, nz, pos, n, weights, wscalar, nc, mesh, flags));
}

}  // extern "C"
