// The binned paint on a workspace the library holds itself: for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  The workspace is one file-scope device
// array per GPU, sized kNoAllocScratchBytes, alive until the process ends.
#include <mutex>

#include "common.hpp"

namespace jpm {

int binned_paint(hipStream_t, int, int, int, const float *, int64_t, const float *, float, int, float *, uint32_t,
                 void *, size_t);
size_t binned_paint_workspace_size(int, int, int, int64_t, int, uint32_t);
bool binned_paint_supported(int, int, int);

constexpr size_t kNoAllocScratchBytes = JPM_NOALLOC_SCRATCH_BYTES;

// Each GPU holds its own copy at its own address.  Host code never passes
// this name to a kernel or a copy: it uses the address that
// hipGetSymbolAddress returns on the current device (scratch_on_device).
static __device__ __attribute__((used)) __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];   // per device ordinal

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and cached.
static int scratch_on_device(void **out) {
  int dev = -1;
  JPM_HIP(hipGetDevice(&dev));
  JPM_REQUIRE(dev >= 0 && dev < kMaxDevices, "device ordinal %d outside 0..%d", dev, kMaxDevices - 1);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    JPM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    JPM_REQUIRE(p != nullptr, "hipGetSymbolAddress returned a null scratch address on device %d", dev);
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return JPM_OK;
}

int noalloc_init_current_device() {
  void *p = nullptr;
  return scratch_on_device(&p);
}

int binned_paint_noalloc(hipStream_t s, int nx, int ny, int nz, const float *pos, int64_t n, const float *weights,
                         float wscalar, int nc, float *mesh, uint32_t flags) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(binned_paint_supported(nx, ny, nz), "mesh (%d,%d,%d) too large for the binned paint", nx, ny, nz);
  JPM_REQUIRE(nc == 1 || nc == 3, "binned paint supports 1 or 3 channels (got %d)", nc);
  JPM_REQUIRE(n >= 0 && n <= kMaxCells, "particle count %lld out of range", (long long)n);
  const size_t need = binned_paint_workspace_size(nx, ny, nz, n, nc, flags);
  JPM_REQUIRE(need <= kNoAllocScratchBytes,
              "the allocation-free paint holds %zu scratch bytes; mesh (%d,%d,%d) with %lld particles and %d "
              "channel(s) needs %zu (use jpm_cic_paint_binned_f32 with a caller workspace)",
              kNoAllocScratchBytes, nx, ny, nz, (long long)n, nc, need);
  void *ws = nullptr;
  rc = scratch_on_device(&ws);
  if (rc) return rc;
  TimingPause no_events;
  return binned_paint(s, nx, ny, nz, pos, n, weights, wscalar, nc, mesh, flags, ws, kNoAllocScratchBytes);
}

}  // namespace jpm
