// Tile geometry shared by the tile sort, the window paint and the window
// gathers (16^3-cell tiles; particles sorted by the tile of their floor cell).
#pragma once

#include <vector>

#include "common.hpp"

namespace jpm {

constexpr int kTile = 16;                       // tile edge in cells
constexpr int kWin = kTile + 1;                 // window edge (CIC +1 corner)
constexpr int kWinCells = kWin * kWin * kWin;   // 4913
constexpr int kChunk = 8192;                    // default max particles per chunk
constexpr int kMaxKeysLds = 32768;              // LDS histogram limit (128 KiB)
constexpr uint32_t kSlice = 8192;               // default particles per count / scatter workgroup
// Accumulation / force-gather workgroup size: their LDS windows (38-39 KiB)
// allow 4 workgroups per CU, so 512 threads keep 32 waves resident.
constexpr int kWindowThreads = 512;

// Sort tuning carried in bits 16..26 of the pipeline flags (0 = default):
// chunk = 2^(9 + c) particles for c = bits 16..19 in 1..6 (1024..32768),
// slice = 2^(11 + s) particles for s = bits 20..23 in 1..7 (4096..262144).
inline uint32_t tuned_chunk(uint32_t flags) {
  const uint32_t c = (flags >> 16) & 0xfu;
  return (c >= 1 && c <= 6) ? (1u << (9 + c)) : uint32_t(kChunk);
}
// Bit 24: the force-VJP gather reads the density adjoint through the caches
// instead of an LDS window.
constexpr uint32_t kTuneRhoGlobal = 1u << 24;
inline bool tuned_rho_lds(uint32_t flags) { return !(flags & kTuneRhoGlobal); }
// Bit 25: XCD slabs.  The dispatcher deals workgroup b to XCD b mod 8; with
// slabs, the chunk (or tile) of workgroup b is (b mod 8) * per + b / 8, so
// each XCD walks one contiguous run of tiles and neighbouring tiles, whose
// windows overlap, share that XCD's L2.
constexpr uint32_t kTuneXcdSlabs = 1u << 25;
inline bool tuned_slabs(uint32_t flags) { return (flags & kTuneXcdSlabs) != 0; }
// Bit 26: the force evaluations' one-channel paint writes its finished
// low-face cells into the mesh, not into a buffer of their own that the
// potential solve's first pass reads (FaceCells); the same bits either way.
constexpr uint32_t kTuneFacesInMesh = 1u << 26;

// Workgroups for `items` work items and the per-XCD run length.
inline unsigned slab_grid(uint32_t items, bool slabs, uint32_t &per) {
  per = (items + 7u) / 8u;
  if (per == 0) per = 1;
  return slabs ? 8u * per : (items ? items : 1u);
}

__device__ __forceinline__ uint32_t slab_item(uint32_t b, uint32_t per, bool slabs) {
  return slabs ? (b & 7u) * per + (b >> 3) : b;
}

// XCD runs for the window gathers (without slabs): of each 8 R consecutive
// workgroups b = 8 R k + 8 q + x (the dispatcher deals b to XCD x), chunk
// 8 R k + R x + q, so each run of R consecutive chunks -- z-neighbouring
// tiles when single-chunk, two y-adjacent z rows of tiles at 256^3 -- runs
// on one XCD at once, and the 64-byte lines that their potential windows'
// rows overlap in (a window row of 21 floats spans three such lines) are
// fetched into its L2 once.  Runs still alternate over the XCDs, which keeps
// clustered tiles spread over them (the slab placement, one contiguous eighth
// of the tiles per XCD, does not).  At 256^3 with 512-thread gathers, against
// dispatch order: force+kick 0.239 -> 0.227 ms, force VJP 0.293 -> 0.280 ms
// (runs of 2: 0.235 / 0.286; of 8: 0.230 / 0.282; of 16: 0.229 / 0.280);
// with the 1024- and 768-thread gathers runs of 32 against 16 and 64:
// 0.2174 / 0.2693 ms against 0.2184 / 0.2703 and 0.2199 / 0.2715.
constexpr uint32_t kXcdRun = 32;
inline unsigned gather_grid(uint32_t items, bool slabs, uint32_t &per) {
  const unsigned g = slab_grid(items, slabs, per);
  return slabs ? g : (g + 8u * kXcdRun - 1u) / (8u * kXcdRun) * (8u * kXcdRun);
}

__device__ __forceinline__ uint32_t gather_item(uint32_t b, uint32_t per, bool slabs) {
  if (slabs) return slab_item(b, per, true);
  const uint32_t blk = b / (8u * kXcdRun), r = b % (8u * kXcdRun);
  return blk * (8u * kXcdRun) + (r & 7u) * kXcdRun + (r >> 3);
}
inline uint32_t tuned_slice(uint32_t flags) {
  const uint32_t v = (flags >> 20) & 0xfu;
  return (v >= 1 && v <= 7) ? (1u << (11 + v)) : kSlice;
}

// A tile's low-face cells (lx = 0, or ly = 0, or lz = 0), numbered in face
// order: the lx = 0 plane (ly * 16 + lz), then ly = 0 with lx > 0
// (256 + (lx - 1) * 16 + lz), then lz = 0 with lx, ly > 0
// (496 + (lx - 1) * 15 + ly - 1) -- 721 cells.
constexpr int kFaces = kTile * kTile + (kTile - 1) * kTile + (kTile - 1) * (kTile - 1);

__host__ __device__ __forceinline__ int face_index(int a, int b, int c) {
  return a == 0 ? b * kTile + c
                : (b == 0 ? kTile * kTile + (a - 1) * kTile + c
                          : kTile * kTile + (kTile - 1) * kTile + (a - 1) * (kTile - 1) + (b - 1));
}

// The low-face cells of a paint kept apart from its mesh.  A tile's face
// cells take the lower neighbours' halos, which the face reduce adds; the
// force paths let it write the finished face values here, contiguous per
// tile, and the potential solve's first pass reads them in place of the
// mesh's (stale) face cells.  Written into the mesh, the lz = 0 face cells
// are single floats a mesh row apart, each write a partial 64-byte line.
// Channel ch, tile T at fb + ch * chan + T * kFaceStride, in face order.
// Needs one simulation whose mesh sides are multiples of 16 (every tile
// full), absolute positions and a paint that replaces the mesh.
constexpr int kFaceStride = 724;   // 721 cells, padded to 16 bytes
struct FaceCells {
  float *fb;            // null: the mesh holds every cell
  uint32_t ny, ty, tz;  // mesh cells along y; tiles along y and z
  size_t chan;          // floats per channel
  FastDiv div_ny;       // row -> (x, y)
};

// A load through the global address space.  For some pointers held in
// structs the compiler otherwise emits a flat load, which also counts
// against the LDS wait counter: every LDS wait after it would wait for the
// memory load too, and a load issued early to hide its latency would not.
__device__ __forceinline__ float4 load_global(const float4 *p) {
  const __attribute__((address_space(1))) float *q = (const __attribute__((address_space(1))) float *)(p);
  return make_float4(q[0], q[1], q[2], q[3]);
}

struct TileGrid {
  int nx, ny, nz;   // mesh
  int tx, ty, tz;   // tiles per axis
  uint32_t K;       // tiles
  FastDiv div_nz, div_ny, div_tz, div_ty;
};

inline TileGrid make_tiles(int nx, int ny, int nz) {
  TileGrid t;
  t.nx = nx;
  t.ny = ny;
  t.nz = nz;
  t.tx = (nx + kTile - 1) / kTile;
  t.ty = (ny + kTile - 1) / kTile;
  t.tz = (nz + kTile - 1) / kTile;
  t.K = uint32_t(t.tx) * t.ty * t.tz;
  t.div_nz = FastDiv(uint32_t(nz));
  t.div_ny = FastDiv(uint32_t(ny));
  t.div_tz = FastDiv(uint32_t(t.tz));
  t.div_ty = FastDiv(uint32_t(t.ty));
  return t;
}

__device__ __forceinline__ int wrapi(int c, int n) {
  // cells already inside [0, n) (nearly all of them) skip the division
  if (uint32_t(c) < uint32_t(n)) return c;
  int r = c % n;
  return r < 0 ? r + n : r;
}

// Floor cell of a particle: absolute -> int32(floor(x)) mod n (painting.py:
// 23, 35-37); relative -> the low corner of enmesh (painting_utils.py:48-57),
// or, when that corner is dropped, the high corner minus one.
__device__ __forceinline__ int floor_cell_abs(float x, int n) { return wrapi(cell_int(floorf(x)), n); }

__device__ __forceinline__ float jmod(float x, float L) {
  float m = fmodf(x, L);
  if (m != 0.0f && m < 0.0f) m = m + L;
  return m;
}

__device__ __forceinline__ int floor_cell_rel(float p, int n) {
  const float L = float(n);
  float q = floorf(jmod(p, L));
  if (q >= 0.0f && q < L) return int(q);
  float q1 = floorf(jmod(p + 1.0f, L));
  if (q1 >= 0.0f && q1 < L) return wrapi(int(q1) - 1, n);
  return 0;  // NaN: every corner is dropped, any tile will do
}

struct ChunkDesc {
  uint32_t tile, start, count, pad;
};

// One simulation as the tile-sort kernels see it.  A batch concatenates its
// simulations: particles [part0, part0 + n) of the positions, tiles
// [tile0, tile0 + t.K) of the tile arrays, cells [cell0, cell0 + M) of the
// meshes, count/scatter workgroups from blk0.
struct SimRef {
  TileGrid t;
  uint32_t tile0, part0, n, cell0, blk0, idx;
};

// A single simulation: the table is one entry held by value.
struct OneSim {
  SimRef s;
  __device__ __forceinline__ const SimRef &tile_sim(uint32_t) const { return s; }
  __device__ __forceinline__ const SimRef &block_sim(uint32_t) const { return s; }
};

// A batch: a device table of B entries, ordered by tile0 and blk0, looked up
// by binary search (the lookups are uniform across a workgroup).
struct SimTable {
  const SimRef *sims;
  int B;
  __device__ __forceinline__ const SimRef &tile_sim(uint32_t tile) const {
    int lo = 0, hi = B - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (sims[mid].tile0 <= tile) lo = mid; else hi = mid - 1;
    }
    return sims[lo];
  }
  __device__ __forceinline__ const SimRef &block_sim(uint32_t blk) const {
    int lo = 0, hi = B - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (sims[mid].blk0 <= blk) lo = mid; else hi = mid - 1;
    }
    return sims[lo];
  }
};

// The result of a tile sort: particles as float4 (x, y, z, global index
// bits) grouped by tile, chunked into runs of at most kChunk particles of one
// tile.  K, n and M are the batch totals.
template <class G>
struct BinsT {
  G g;
  uint32_t K, n, M, maxchunks, chunk;
  bool rel, slabs;
  const float4 *spos;
  const ChunkDesc *desc;
  const uint32_t *nchunk, *choff, *total_chunks;
  // set by a paint of these bins that met a NaN or +-Inf deposit; cleared by
  // the sort
  uint32_t *nonfinite;
};
using Bins = BinsT<OneSim>;
using BatchBins = BinsT<SimTable>;

// A batch of simulations: the host description and the device table.
// Simulations of one mesh shape are contiguous (a "class"), so each class's
// meshes and half spectra form one strided batch for hipFFT.
struct SimClass {
  int nx, ny, nz, first, count;
  uint32_t cell0, half0;
};

struct SimBatch {
  int B = 0;
  bool rel = false;
  uint32_t K = 0, N = 0, M = 0, H = 0, G = 0, maxK = 0, chunk = kChunk, slice = kSlice;
  bool rho_lds = true, slabs = false;
  std::vector<SimRef> host;
  std::vector<SimClass> classes;
  SimRef *dev = nullptr;
};

inline SimRef one_sim(int nx, int ny, int nz, int64_t n) {
  SimRef r;
  r.t = make_tiles(nx, ny, nz);
  r.tile0 = r.part0 = r.cell0 = r.blk0 = r.idx = 0;
  r.n = uint32_t(n);
  return r;
}

int bin_sort(hipStream_t s, int nx, int ny, int nz, const float *pos, int64_t n, uint32_t flags, void *ws,
             size_t ws_bytes, Bins &out, const uint32_t *counts_in = nullptr);
size_t bin_sort_workspace_size(int nx, int ny, int nz, int64_t n, uint32_t flags);
Bins bins_view(int nx, int ny, int nz, int64_t n, uint32_t flags, const void *ws);
template <class G>
int bins_order(hipStream_t s, const BinsT<G> &b, uint32_t *order);
size_t batch_sort_workspace_size(const SimBatch &bt);
size_t batch_paint_workspace_size(const SimBatch &bt, int nc);
int batch_sort(hipStream_t s, const SimBatch &bt, const float *pos, void *ws, size_t ws_bytes, BatchBins &out,
               const uint32_t *counts_in = nullptr);
int zero_u32(hipStream_t s, uint32_t *p, uint32_t n);
int bins_paint(hipStream_t s, const Bins &b, const float *weight, float wscalar, int nc, float *mesh, uint32_t flags,
               void *ws, size_t ws_bytes, const float *wmax_in = nullptr, const FaceCells *faces = nullptr);
// The potential solves (kspace.hip); faces: the painted mesh's face cells
// kept apart, on the shapes potential_faces_ok accepts, for a one-mesh solve.
bool potential_faces_ok(int nx, int ny, int nz);
int potential_solve(hipStream_t s, int nx, int ny, int nz, int batch, const float *in, void *spec, void *spec2,
                    float *out, float r_split, float scale, const float *scale_dev, uint32_t flags,
                    const FaceCells *faces = nullptr);
int potential_solve_div(hipStream_t s, int nx, int ny, int nz, int batch, const float *g3, size_t chan, float *rho,
                        void *spec, void *spec2, float *out, float r_split, float scale, const float *scale_dev,
                        uint32_t flags);
BatchBins batch_view(const SimBatch &bt, const void *ws);
int batch_paint(hipStream_t s, const SimBatch &bt, const BatchBins &b, const float *weight, int nc, float *mesh,
                void *ws, size_t ws_bytes, const float *wmax_in = nullptr);

// One corner axis of a particle relative to its tile origin: local window
// index (or -1 when the corner is dropped / outside the window) and weight.
struct LAxis {
  int l0, l1;
  float w0, w1;
};

__device__ __forceinline__ LAxis laxis_abs(float x, int origin, int n) {
  LAxis a;
  const float f = floorf(x);
  const float f1 = f + 1.0f;
  a.w0 = 1.0f - fabsf(x - f);
  a.w1 = 1.0f - fabsf(x - f1);
  const int c0 = wrapi(cell_int(f), n);
  a.l0 = c0 - origin;  // the floor cell lies in the tile by construction
  a.l1 = a.l0 + 1;
  return a;
}

// enmesh corner (painting_utils.py:48-69) mapped into the window; corners
// that land outside [origin, origin + 16] (rounding edge cases) report -1 and
// are deposited by the caller through the overflow path.
__device__ __forceinline__ void lcorner_rel(float p, float o, int n, int origin, int &l, float &w, int &cell) {
  const float L = float(n);
  float q = floorf(jmod(p + o, L));
  float nd = p - q;
  nd = nd - rintf(nd / L) * L;
  const bool ok = (q >= 0.0f) && (q < L);
  cell = ok ? int(q) : -1;
  w = ok ? 1.0f - fabsf(nd) : 0.0f;
  if (!ok) {
    l = -2;  // dropped: no deposit at all
    return;
  }
  int d = int(q) - origin;
  if (d < 0) d += n;  // window may wrap past the mesh end
  l = (d >= 0 && d < kWin) ? d : -1;
}

__device__ __forceinline__ int tile_len(int it, int ntile, int n) {
  return (it == ntile - 1) ? n - it * kTile : kTile;
}


// Tile origin of a tile index.
__device__ __forceinline__ void tile_origin(uint32_t tile, const TileGrid &t, int &itx, int &ity, int &itz) {
  const uint32_t txy = t.div_tz.div(tile);
  itz = int(tile - txy * uint32_t(t.tz));
  itx = int(t.div_ty.div(txy));
  ity = int(txy - uint32_t(itx) * uint32_t(t.ty));
}

}  // namespace jpm
