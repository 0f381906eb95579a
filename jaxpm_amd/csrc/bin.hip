// Binned CIC paint: counting sort of particles into 16^3-cell tiles, LDS
// accumulation of each chunk of a tile's particles into the tile's (17^3)
// window, and a gather-reduce that sums the windows covering each cell.
// No global float atomics.  The window sums are 64-bit fixed point (on gfx950
// ds_add_u64 is several times faster than ds_add_f32), so the result is
// bitwise reproducible whatever the order of deposits.
//
// Passes (all stream-ordered, no host synchronisation):
//   count    per-workgroup LDS histogram of tile keys (lanes read consecutive
//            particles; each run of equal keys within a wavefront adds one
//            count -- Lagrangian order keeps runs long), then one global
//            atomic per non-empty (tile, workgroup) pair
//   scan     tile start = exclusive scan of the counts; chunk offsets
//            (ceil(count / CH) chunks per tile)
//   chunks   chunk descriptors (tile, start, count)
//   scatter  each workgroup recounts its slice in LDS, reserves one range per
//            non-empty tile with a global atomic on the tile's cursor, and
//            writes its particles there as float4 (x, y, z, index bits)
//   accum    one workgroup per (chunk, channel): LDS window of int64, the
//            deposits, window -> partial
//   reduce   one workgroup per tile: for each of its cells, the tile's own
//            chunk windows plus the +1 halos of the (up to 7) lower
//            neighbour tiles, summed in a fixed order and converted
#include <type_traits>

#include "tiles.hpp"

namespace jpm {

constexpr int kBinThreads = 256;                // hist / scatter workgroup size
constexpr int kScanItems = 8;                   // scan: elements per thread
constexpr int kScanBlock = 256 * kScanItems;

// Floor cell of particle p (global index; `local` its index within its
// simulation, which relative mode decodes into the lattice node).
template <bool REL>
__device__ __forceinline__ void floor_cells_xyz(float x, float y, float z, uint32_t local, const TileGrid &t, int &cx,
                                                int &cy, int &cz) {
  if (REL) {
    uint32_t r = t.div_nz.div(local);
    uint32_t k = local - r * uint32_t(t.nz);
    uint32_t i = t.div_ny.div(r);
    uint32_t j = r - i * uint32_t(t.ny);
    cx = floor_cell_rel(float(i) + x, t.nx);
    cy = floor_cell_rel(float(j) + y, t.ny);
    cz = floor_cell_rel(float(k) + z, t.nz);
  } else {
    cx = floor_cell_abs(x, t.nx);
    cy = floor_cell_abs(y, t.ny);
    cz = floor_cell_abs(z, t.nz);
  }
}

template <bool REL>
__device__ __forceinline__ void floor_cells(const float *__restrict__ pos, uint32_t p, uint32_t local,
                                            const TileGrid &t, int &cx, int &cy, int &cz) {
  floor_cells_xyz<REL>(pos[3 * size_t(p)], pos[3 * size_t(p) + 1], pos[3 * size_t(p) + 2], local, t, cx, cy, cz);
}

__device__ __forceinline__ uint32_t tile_key(int cx, int cy, int cz, const TileGrid &t) {
  return (uint32_t(cx / kTile) * uint32_t(t.ty) + uint32_t(cy / kTile)) * uint32_t(t.tz) + uint32_t(cz / kTile);
}

// --------------------------------------------------------------- count

// Runs of equal keys among the 64 lanes of a wavefront: `heads` has bit l set
// when lane l starts a run (its key differs from lane l-1's, or l is lane 0);
// a head's run ends where the next head begins or at the last active lane
// (active lanes are always a prefix of the wavefront).
__device__ __forceinline__ unsigned long long run_heads(uint32_t key, bool active) {
  const int lane = threadIdx.x & 63;
  const uint32_t prev = __shfl_up(key, 1);
  const bool head = active && (lane == 0 || key != prev);
  return __ballot(head);
}

__device__ __forceinline__ uint32_t run_length(unsigned long long heads, unsigned long long act) {
  const int lane = threadIdx.x & 63;
  const unsigned long long above = heads & ~((2ull << lane) - 1ull);
  const int next = above ? __ffsll((long long)above) - 1 : 64;
  const int last_active = 63 - __clzll((long long)act);
  return uint32_t(min(next, last_active + 1) - lane);
}

// The particle range [p0, p1) of count/scatter workgroup `blk` of simulation s.
__device__ __forceinline__ void block_range(const SimRef &s, uint32_t blk, uint32_t slice, uint32_t &p0,
                                            uint32_t &p1) {
  p0 = s.part0 + (blk - s.blk0) * slice;
  p1 = min(s.part0 + s.n, p0 + slice);
}

// LDS histogram of the (simulation-local) tile keys of particles [p0, p1).
template <bool REL>
__device__ __forceinline__ void slice_histogram(const float *__restrict__ pos, uint32_t p0, uint32_t p1,
                                                const SimRef &s, uint32_t *sh) {
  for (uint32_t b0 = p0; b0 < p1; b0 += blockDim.x) {
    const uint32_t p = b0 + threadIdx.x;
    const bool active = p < p1;
    uint32_t key = 0xffffffffu;
    if (active) {
      int cx, cy, cz;
      floor_cells<REL>(pos, p, p - s.part0, s.t, cx, cy, cz);
      key = tile_key(cx, cy, cz, s.t);
    }
    const unsigned long long act = __ballot(active);
    const unsigned long long heads = run_heads(key, active);
    if (active && ((heads >> (threadIdx.x & 63)) & 1ull)) atomicAdd(&sh[key], run_length(heads, act));
  }
}

template <bool REL, class G>
__global__ void __launch_bounds__(kBinThreads)
bin_count_kernel(const float *__restrict__ pos, uint32_t slice, G g, uint32_t *__restrict__ count) {
  extern __shared__ uint32_t sh[];
  const SimRef &s = g.block_sim(blockIdx.x);
  const uint32_t K = s.t.K, tile0 = s.tile0;
  for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) sh[k] = 0;
  __syncthreads();
  uint32_t p0, p1;
  block_range(s, blockIdx.x, slice, p0, p1);
  slice_histogram<REL>(pos, p0, p1, s, sh);
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < K; k += blockDim.x)
    if (sh[k]) atomicAdd(&count[tile0 + k], sh[k]);
}

// Tile counts for meshes of more than kMaxKeysLds tiles, where no LDS
// histogram fits: each run of equal keys in a wavefront adds its length with
// one global atomic (runs along the particle order keep these few).
template <bool REL, class G>
__global__ void __launch_bounds__(kBinThreads)
bin_count_global_kernel(const float *__restrict__ pos, uint32_t slice, G g, uint32_t *__restrict__ count) {
  const SimRef &s = g.block_sim(blockIdx.x);
  uint32_t p0, p1;
  block_range(s, blockIdx.x, slice, p0, p1);
  for (uint32_t b0 = p0; b0 < p1; b0 += blockDim.x) {
    const uint32_t p = b0 + threadIdx.x;
    const bool active = p < p1;
    uint32_t key = 0xffffffffu;
    if (active) {
      int cx, cy, cz;
      floor_cells<REL>(pos, p, p - s.part0, s.t, cx, cy, cz);
      key = tile_key(cx, cy, cz, s.t);
    }
    const unsigned long long act = __ballot(active);
    const unsigned long long heads = run_heads(key, active);
    if (active && ((heads >> (threadIdx.x & 63)) & 1ull)) atomicAdd(&count[s.tile0 + key], run_length(heads, act));
  }
}

// ---------------------------------------------------------------- scan

// In-place exclusive scan of blocks of kScanBlock elements; block totals to
// sums[blockIdx.x].
__global__ void __launch_bounds__(256) scan_blocks_kernel(uint32_t *__restrict__ a, uint32_t n,
                                                          uint32_t *__restrict__ sums) {
  __shared__ uint32_t sh[256];
  const size_t base = size_t(blockIdx.x) * kScanBlock + size_t(threadIdx.x) * kScanItems;
  uint32_t v[kScanItems];
  uint32_t tot = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = (base + i < n) ? a[base + i] : 0u;
    tot += v[i];
  }
  sh[threadIdx.x] = tot;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    uint32_t x = threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
    __syncthreads();
    sh[threadIdx.x] += x;
    __syncthreads();
  }
  uint32_t run = sh[threadIdx.x] - tot;  // exclusive prefix of this thread
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) a[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == 255) sums[blockIdx.x] = sh[255];
}

// Single-workgroup exclusive scan of `sums` (any length); the grand total
// goes to *total.
__global__ void __launch_bounds__(256) scan_sums_kernel(uint32_t *__restrict__ sums, uint32_t nb,
                                                        uint32_t *__restrict__ total) {
  __shared__ uint32_t sh[256];
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < nb; c0 += 256) {
    const uint32_t i = c0 + threadIdx.x;
    const uint32_t v = i < nb ? sums[i] : 0u;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      uint32_t x = threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
      __syncthreads();
      sh[threadIdx.x] += x;
      __syncthreads();
    }
    if (i < nb) sums[i] = carry + sh[threadIdx.x] - v;
    const uint32_t chunk_total = sh[255];
    __syncthreads();
    carry += chunk_total;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void __launch_bounds__(256) zero_u32_kernel(uint32_t *__restrict__ a, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = 0u;
}

__global__ void __launch_bounds__(256) scan_add_kernel(uint32_t *__restrict__ a, uint32_t n,
                                                       const uint32_t *__restrict__ sums) {
  const size_t base = size_t(blockIdx.x) * kScanBlock + size_t(threadIdx.x) * kScanItems;
  const uint32_t add = sums[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) a[base + i] += add;
}

static int exclusive_scan(hipStream_t s, uint32_t *a, uint32_t n, uint32_t *sums, uint32_t *total) {
  const uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
  hipLaunchKernelGGL(scan_blocks_kernel, dim3(nb), dim3(256), 0, s, a, n, sums);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(256), 0, s, sums, nb, total);
  hipLaunchKernelGGL(scan_add_kernel, dim3(nb), dim3(256), 0, s, a, n, sums);
  JPM_LAUNCHED("exclusive_scan");
  return JPM_OK;
}

// nchunk[k] = ceil(count[k] / kChunk).
__global__ void __launch_bounds__(256)
tile_chunks_kernel(const uint32_t *__restrict__ count, uint32_t K, uint32_t chunk, uint32_t *__restrict__ nchunk) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < K; k += gridDim.x * blockDim.x)
    nchunk[k] = (count[k] + chunk - 1) / chunk;
}

__global__ void __launch_bounds__(256)
chunk_desc_kernel(const uint32_t *__restrict__ tile_start, const uint32_t *__restrict__ tile_count,
                  const uint32_t *__restrict__ chunk_off, uint32_t K, uint32_t chunk, ChunkDesc *__restrict__ desc) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < K; k += gridDim.x * blockDim.x) {
    const uint32_t st = tile_start[k], cnt = tile_count[k];
    uint32_t c = chunk_off[k];
    for (uint32_t o = 0; o < cnt; o += chunk, ++c) {
      ChunkDesc d;
      d.tile = k;
      d.start = st + o;
      d.count = min(chunk, cnt - o);
      d.pad = cnt > chunk ? 1u : 0u;  // tile split over several chunks
      desc[c] = d;
    }
  }
}

// The whole tile table of a sort of at most kTableMaxK tiles in one launch
// of ceil(K / 1024) workgroups: tile starts and scatter cursors (exclusive
// scan of the counts), chunks per tile and their offsets (exclusive scan),
// the chunk descriptors, and the totals [particles, chunks].  Each thread
// owns one tile; each workgroup sums the counts before its range itself and
// scans its own by wavefront shuffles.  The same values as count -> scans ->
// tile_chunks -> chunk_desc, which larger tile grids use.
constexpr int kTableThreads = 1024;
constexpr uint32_t kTableMaxK = 32768;

__global__ void __launch_bounds__(kTableThreads)
tile_table_kernel(const uint32_t *__restrict__ count, uint32_t K, uint32_t chunk, uint32_t *__restrict__ tstart,
                  uint32_t *__restrict__ cursor, uint32_t *__restrict__ nchunk, uint32_t *__restrict__ choff,
                  uint32_t *__restrict__ totals, ChunkDesc *__restrict__ desc) {
  // workgroup b owns tiles [b T, (b + 1) T), one per thread, T = kTableThreads
  __shared__ uint32_t sp[kTableThreads / 64], sc[kTableThreads / 64];
  const uint32_t t = threadIdx.x, lane = t & 63u, wid = t >> 6;
  const uint32_t first = blockIdx.x * kTableThreads, k = first + t;
  // particles and chunks of the tiles before this workgroup's (summed by
  // every workgroup for itself: at most K counts, read from L2)
  uint32_t bp = 0, bc = 0;
  for (uint32_t j = t; j < first; j += kTableThreads) {
    const uint32_t c = count[j];
    bp += c;
    bc += (c + chunk - 1) / chunk;
  }
  const uint32_t np = k < K ? count[k] : 0u, nc = (np + chunk - 1) / chunk;
  // one pass of wavefront shuffles: the base sums reduced, this
  // workgroup's (np, nc) scanned inclusively
  uint32_t xp = np, xc = nc;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t yp = __shfl_up(xp, off), yc = __shfl_up(xc, off);
    if (lane >= uint32_t(off)) {
      xp += yp;
      xc += yc;
    }
    bp += __shfl_xor(bp, off);
    bc += __shfl_xor(bc, off);
  }
  __shared__ uint32_t bsp[kTableThreads / 64], bsc[kTableThreads / 64];
  if (lane == 63u) {
    sp[wid] = xp;
    sc[wid] = xc;
    bsp[wid] = bp;
    bsc[wid] = bc;
  }
  __syncthreads();
  uint32_t basep = 0, basec = 0;
  for (uint32_t w = 0; w < kTableThreads / 64; ++w) {
    basep += bsp[w];
    basec += bsc[w];
    if (w < wid) {
      xp += sp[w];
      xc += sc[w];
    }
  }
  const uint32_t pb = basep + xp - np;
  uint32_t cb = basec + xc - nc;
  if (k == K - 1) {
    totals[0] = pb + np;
    totals[1] = cb + nc;
    totals[2] = 0u;   // the non-finite flag of the paints of this sort
  }
  if (k < K) {
    tstart[k] = pb;
    cursor[k] = pb;
    nchunk[k] = nc;
    choff[k] = cb;
    for (uint32_t o = 0; o < np; o += chunk, ++cb) {
      ChunkDesc d;
      d.tile = k;
      d.start = pb + o;
      d.count = min(chunk, np - o);
      d.pad = np > chunk ? 1u : 0u;
      desc[cb] = d;
    }
  }
}

// -------------------------------------------------------------- scatter

template <bool REL, class G>
__global__ void __launch_bounds__(kBinThreads)
bin_scatter_kernel(const float *__restrict__ pos, uint32_t slice, G g, uint32_t *__restrict__ cursor,
                   float4 *__restrict__ spos) {
  extern __shared__ uint32_t cur[];
  const SimRef &s = g.block_sim(blockIdx.x);
  const uint32_t K = s.t.K, tile0 = s.tile0;
  for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) cur[k] = 0;
  __syncthreads();
  uint32_t p0, p1;
  block_range(s, blockIdx.x, slice, p0, p1);
  slice_histogram<REL>(pos, p0, p1, s, cur);
  __syncthreads();
  // reserve this workgroup's range in every tile it touches
  for (uint32_t k = threadIdx.x; k < K; k += blockDim.x)
    if (cur[k]) cur[k] = atomicAdd(&cursor[tile0 + k], cur[k]);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  for (uint32_t b0 = p0; b0 < p1; b0 += blockDim.x) {
    const uint32_t p = b0 + threadIdx.x;
    const bool active = p < p1;
    uint32_t key = 0xffffffffu;
    float x = 0.0f, y = 0.0f, z = 0.0f;
    if (active) {
      x = pos[3 * size_t(p)];
      y = pos[3 * size_t(p) + 1];
      z = pos[3 * size_t(p) + 2];
      int cx, cy, cz;
      floor_cells<REL>(pos, p, p - s.part0, s.t, cx, cy, cz);
      key = tile_key(cx, cy, cz, s.t);
    }
    const unsigned long long act = __ballot(active);
    const unsigned long long heads = run_heads(key, active);
    const bool head = active && ((heads >> lane) & 1ull);
    uint32_t base = 0;
    if (head) base = atomicAdd(&cur[key], run_length(heads, act));
    // each lane takes the base of the run it belongs to (its nearest head at
    // or below it) plus its distance from that head
    const unsigned long long below = heads & ((2ull << lane) - 1ull);
    const int my_head = below ? 63 - __clzll((long long)below) : 0;
    base = __shfl(base, my_head);
    if (active) spos[base + uint32_t(lane - my_head)] = make_float4(x, y, z, __uint_as_float(p));
  }
}

// bin_scatter_kernel with the slice held in registers: THREADS threads, PER
// particles each (slice = THREADS PER), loaded and keyed once for both the
// histogram and the scatter, so the positions are read once.  1024 x 8 and
// 1024 x 16 hold one workgroup per CU (their registers); 512 x 8, three.
constexpr int kScatterRegThreads = 1024;

template <bool REL, class G, int PER, int THREADS>
__global__ void __launch_bounds__(THREADS)
bin_scatter_reg_kernel(const float *__restrict__ pos, uint32_t slice, G g, uint32_t *__restrict__ cursor,
                       float4 *__restrict__ spos) {
  extern __shared__ uint32_t cur[];
  const SimRef &s = g.block_sim(blockIdx.x);
  const uint32_t K = s.t.K, tile0 = s.tile0;
  for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) cur[k] = 0;
  uint32_t p0, p1;
  block_range(s, blockIdx.x, slice, p0, p1);
  const int lane = threadIdx.x & 63;
  float x[PER], y[PER], z[PER];
  uint32_t key[PER];
#pragma unroll
  for (int it = 0; it < PER; ++it) {
    const uint32_t p = p0 + uint32_t(it) * THREADS + threadIdx.x;
    key[it] = 0xffffffffu;
    x[it] = y[it] = z[it] = 0.0f;
    if (p < p1) {
      x[it] = pos[3 * size_t(p)];
      y[it] = pos[3 * size_t(p) + 1];
      z[it] = pos[3 * size_t(p) + 2];
    }
  }
#pragma unroll
  for (int it = 0; it < PER; ++it) {
    const uint32_t p = p0 + uint32_t(it) * THREADS + threadIdx.x;
    if (p < p1) {
      int cx, cy, cz;
      floor_cells_xyz<REL>(x[it], y[it], z[it], p - s.part0, s.t, cx, cy, cz);
      key[it] = tile_key(cx, cy, cz, s.t);
    }
  }
  __syncthreads();
  // LDS histogram of the slice, one atomic per run of equal keys in a wavefront
#pragma unroll
  for (int it = 0; it < PER; ++it) {
    const uint32_t p = p0 + uint32_t(it) * THREADS + threadIdx.x;
    const bool active = p < p1;
    const unsigned long long act = __ballot(active);
    const unsigned long long heads = run_heads(key[it], active);
    if (active && ((heads >> lane) & 1ull)) atomicAdd(&cur[key[it]], run_length(heads, act));
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < K; k += blockDim.x)
    if (cur[k]) cur[k] = atomicAdd(&cursor[tile0 + k], cur[k]);
  __syncthreads();
#pragma unroll
  for (int it = 0; it < PER; ++it) {
    const uint32_t p = p0 + uint32_t(it) * THREADS + threadIdx.x;
    const bool active = p < p1;
    const unsigned long long act = __ballot(active);
    const unsigned long long heads = run_heads(key[it], active);
    const bool head = active && ((heads >> lane) & 1ull);
    uint32_t base = 0;
    if (head) base = atomicAdd(&cur[key[it]], run_length(heads, act));
    const unsigned long long below = heads & ((2ull << lane) - 1ull);
    const int my_head = below ? 63 - __clzll((long long)below) : 0;
    base = __shfl(base, my_head);
    if (active) spos[base + uint32_t(lane - my_head)] = make_float4(x[it], y[it], z[it], __uint_as_float(p));
  }
}

// Scatter for meshes of more than kMaxKeysLds tiles: each wavefront run
// reserves its range with one global atomic on the tile's cursor.
template <bool REL, class G>
__global__ void __launch_bounds__(kBinThreads)
bin_scatter_global_kernel(const float *__restrict__ pos, uint32_t slice, G g, uint32_t *__restrict__ cursor,
                          float4 *__restrict__ spos) {
  const SimRef &s = g.block_sim(blockIdx.x);
  uint32_t p0, p1;
  block_range(s, blockIdx.x, slice, p0, p1);
  const int lane = threadIdx.x & 63;
  for (uint32_t b0 = p0; b0 < p1; b0 += blockDim.x) {
    const uint32_t p = b0 + threadIdx.x;
    const bool active = p < p1;
    uint32_t key = 0xffffffffu;
    float x = 0.0f, y = 0.0f, z = 0.0f;
    if (active) {
      x = pos[3 * size_t(p)];
      y = pos[3 * size_t(p) + 1];
      z = pos[3 * size_t(p) + 2];
      int cx, cy, cz;
      floor_cells<REL>(pos, p, p - s.part0, s.t, cx, cy, cz);
      key = tile_key(cx, cy, cz, s.t);
    }
    const unsigned long long act = __ballot(active);
    const unsigned long long heads = run_heads(key, active);
    const bool head = active && ((heads >> lane) & 1ull);
    uint32_t base = 0;
    if (head) base = atomicAdd(&cursor[s.tile0 + key], run_length(heads, act));
    const unsigned long long below = heads & ((2ull << lane) - 1ull);
    const int my_head = below ? 63 - __clzll((long long)below) : 0;
    base = __shfl(base, my_head);
    if (active) spos[base + uint32_t(lane - my_head)] = make_float4(x, y, z, __uint_as_float(p));
  }
}

// ----------------------------------------------------------------- accum

// The +1 halo of a full (16^3) single-chunk tile in absolute mode is stored
// compactly: the lx = 16 plane (17 x 17, index ly * 17 + lz), then the ly = 16
// cells with lx < 16 (16 x 17), then the lz = 16 cells with lx, ly < 16
// (16 x 16) -- 817 floats.
constexpr int kH0 = kWin * kWin, kH1 = kH0 + kTile * kWin, kH = kH1 + kTile * kTile;

__device__ __forceinline__ int halo_index(int lx, int ly, int lz) {
  return lx == kTile ? ly * kWin + lz : (ly == kTile ? kH0 + lx * kWin + lz : kH1 + lx * kTile + ly);
}

// A tile's low-face cells (lx = 0, or ly = 0, or lz = 0) in the order the
// reduce visits them: the lx = 0 plane (ly * 16 + lz), then ly = 0 with
// lx > 0, then lz = 0 with lx, ly > 0 -- 721 cells.  Their own values follow
// the halo in the tile's slot (the same floats the mesh receives), so the
// reduce reads them there, contiguous, instead of from the mesh, where the
// lz = 0 cells lie single floats a mesh row apart (their reads cost a
// 64-byte line each: 0.064 -> 0.046 ms for one channel at 256^3).
static_assert(kH + kFaces <= 2 * kWinCells, "halo and own faces fit a window slot (kWinCells int64)");

__device__ __forceinline__ void face_cell(int e, int &a, int &b, int &c) {
  if (e < kTile * kTile) {
    a = 0; b = e >> 4; c = e & 15;
  } else if (e < kTile * kTile + (kTile - 1) * kTile) {
    const int f = e - kTile * kTile;
    a = 1 + (f >> 4); b = 0; c = f & 15;
  } else {
    const int f = e - kTile * kTile - (kTile - 1) * kTile;
    a = 1 + f / (kTile - 1); b = 1 + f % (kTile - 1); c = 0;
  }
}

// The window corners of a sorted particle (record rec) of a tile with
// origin (ox, oy, oz): per axis the local window index of its two corners
// (relative mode: -2 when enmesh drops the corner, -1 when rounding put it
// outside the window), their weights and (relative mode) global cells.
template <bool REL>
__device__ __forceinline__ void window_corners(const float4 &rec, const TileGrid &t, uint32_t part0, int ox, int oy,
                                               int oz, int lx[2], int ly[2], int lz[2], int gx[2], int gy[2],
                                               int gz[2], float wx[2], float wy[2], float wz[2]) {
  const float x = rec.x, y = rec.y, z = rec.z;
  if (REL) {
    const uint32_t local = __float_as_uint(rec.w) - part0;
    uint32_t r = t.div_nz.div(local);
    uint32_t k = local - r * uint32_t(t.nz);
    uint32_t i = t.div_ny.div(r);
    uint32_t j = r - i * uint32_t(t.ny);
    const float px = float(i) + x, py = float(j) + y, pz = float(k) + z;
    lcorner_rel(px, 0.0f, t.nx, ox, lx[0], wx[0], gx[0]);
    lcorner_rel(px, 1.0f, t.nx, ox, lx[1], wx[1], gx[1]);
    lcorner_rel(py, 0.0f, t.ny, oy, ly[0], wy[0], gy[0]);
    lcorner_rel(py, 1.0f, t.ny, oy, ly[1], wy[1], gy[1]);
    lcorner_rel(pz, 0.0f, t.nz, oz, lz[0], wz[0], gz[0]);
    lcorner_rel(pz, 1.0f, t.nz, oz, lz[1], wz[1], gz[1]);
  } else {
    LAxis ax = laxis_abs(x, ox, t.nx), ay = laxis_abs(y, oy, t.ny), az = laxis_abs(z, oz, t.nz);
    lx[0] = ax.l0; lx[1] = ax.l1; wx[0] = ax.w0; wx[1] = ax.w1;
    ly[0] = ay.l0; ly[1] = ay.l1; wy[0] = ay.w0; wy[1] = ay.w1;
    lz[0] = az.l0; lz[1] = az.l1; wz[0] = az.w0; wz[1] = az.w1;
  }
}

// Stores one channel's finished window.  A tile with a single chunk stores
// its window rounded once to float.  A tile split over several chunks keeps
// the exact int64 windows: which particle lands in which chunk depends on the
// scatter's atomic order, so only exact sums across its chunks are
// reproducible bit for bit.  slot: the window's index in the partial buffer;
// dst: the channel's mesh at the simulation's first cell.
template <bool REL>
__device__ __forceinline__ void write_window(const unsigned long long *win, const ChunkDesc &d, const TileGrid &t,
                                             int itx, int ity, int itz, double scale, size_t slot,
                                             float *__restrict__ partial, float *__restrict__ dst, int accumulate) {
  const int ox = itx * kTile, oy = ity * kTile, oz = itz * kTile;
  if (!REL && d.pad == 0) {
    // Absolute positions never leave the window, so the cells of a
    // single-chunk tile hold their own deposits complete: they go straight
    // to the mesh, and only the +1 halo (the low faces of the upper
    // neighbours) is kept for the reduce.
    const double inv = 1.0 / scale;
    const int Lx = tile_len(itx, t.tx, t.nx), Ly = tile_len(ity, t.ty, t.ny), Lz = tile_len(itz, t.tz, t.nz);
    float *out = partial + slot * 2;
    if (Lx == kTile && Ly == kTile && Lz == kTile) {
      // interior: each 16 lanes write one 64-byte row of the mesh
      for (int e = threadIdx.x; e < kTile * kTile * kTile; e += blockDim.x) {
        const int lx = e >> 8, ly = (e >> 4) & 15, lz = e & 15;
        const float v = float(double((long long)win[(lx * kWin + ly) * kWin + lz]) * inv);
        float *cellp = dst + (uint32_t(ox + lx) * uint32_t(t.ny) + uint32_t(oy + ly)) * uint32_t(t.nz) + uint32_t(oz + lz);
        *cellp = accumulate ? *cellp + v : v;
      }
      // halo, compact (halo_index): the lx = 16 plane, then ly = 16 (lx <
      // 16), then lz = 16 (lx, ly < 16), so both this write and the reduce's
      // reads of each face run along consecutive floats
      for (int e = threadIdx.x; e < kH; e += blockDim.x) {
        int lx, ly, lz;
        if (e < kH0) {
          lx = kTile; ly = e / kWin; lz = e % kWin;
        } else if (e < kH1) {
          const int f = e - kH0;
          lx = f / kWin; ly = kTile; lz = f % kWin;
        } else {
          const int f = e - kH1;
          lx = f >> 4; ly = f & 15; lz = kTile;
        }
        out[e] = float(double((long long)win[(lx * kWin + ly) * kWin + lz]) * inv);
      }
      // own face cells (face_cell order), the same floats the mesh received
      if (!accumulate)
        for (int e = threadIdx.x; e < kFaces; e += blockDim.x) {
          int a, b, c;
          face_cell(e, a, b, c);
          out[kH + e] = float(double((long long)win[(a * kWin + b) * kWin + c]) * inv);
        }
    } else {
      for (int i = threadIdx.x; i < kWinCells; i += blockDim.x) {
        const int lz = i % kWin, lxy = i / kWin, ly = lxy % kWin, lx = lxy / kWin;
        const float v = float(double((long long)win[i]) * inv);
        if (lx < Lx && ly < Ly && lz < Lz) {
          float *cellp = dst + (uint32_t(ox + lx) * uint32_t(t.ny) + uint32_t(oy + ly)) * uint32_t(t.nz) + uint32_t(oz + lz);
          *cellp = accumulate ? *cellp + v : v;
        } else {
          out[i] = v;
        }
      }
    }
  } else if (d.pad == 0) {
    const double inv = 1.0 / scale;
    float *out = partial + slot * 2;
    for (int i = threadIdx.x; i < kWinCells; i += blockDim.x) out[i] = float(double((long long)win[i]) * inv);
  } else {
    unsigned long long *out = reinterpret_cast<unsigned long long *>(partial) + slot;
    for (int i = threadIdx.x; i < kWinCells; i += blockDim.x) out[i] = win[i];
  }
}

// A deposit that is NaN or +-Inf (a non-finite weight or position) does not
// enter the fixed-point window: it sets *nonfinite, and bin_reduce_kernel
// adds it to the finished mesh as a float, so it reaches exactly the cells
// it touches, as in a float scatter-add.
template <bool REL, class G>
__global__ void __launch_bounds__(kWindowThreads)
bin_accum_kernel(const float4 *__restrict__ spos, const float *__restrict__ weight, int nc, float wscalar,
                 const ChunkDesc *__restrict__ desc, const uint32_t *__restrict__ nchunks_total, G g, uint32_t M,
                 const float *__restrict__ wmax_dev, float wmax_host, uint32_t n, float *__restrict__ partial,
                 float *__restrict__ overflow, int accumulate, float *__restrict__ mesh, uint32_t per, bool slabs,
                 uint32_t *__restrict__ nonfinite) {
  __shared__ unsigned long long win[kWinCells];
  // (chunk, channel) of this workgroup: the nc channels of one chunk go to
  // one XCD back to back (workgroups b, b + 8, ...), so their reads of the
  // chunk's particles and weights share that XCD's L2
  const uint32_t xcd = blockIdx.x & 7u, run = blockIdx.x >> 3;
  const uint32_t c = slabs ? xcd * per + run / uint32_t(nc) : (run / uint32_t(nc)) * 8u + xcd;
  const int ch = int(run % uint32_t(nc));
  if (c >= *nchunks_total) return;
  const ChunkDesc d = desc[c];
  const SimRef &sim = g.tile_sim(d.tile);
  const TileGrid &t = sim.t;
  const uint32_t part0 = sim.part0, cell0 = sim.cell0;
  int itx, ity, itz;
  tile_origin(d.tile - sim.tile0, t, itx, ity, itz);
  const int ox = itx * kTile, oy = ity * kTile, oz = itz * kTile;
  // Software pipeline over this thread's particles q, q + B, ...: particle
  // q deposits while the weight of q + B (whose record arrived an iteration
  // ago) and the record of q + 2B load.
  const uint32_t cnt = d.count, B = blockDim.x;
  const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float4 rec1 = threadIdx.x < cnt ? load_global(spos + (d.start + threadIdx.x)) : zero4;
  float4 rec2 = threadIdx.x + B < cnt ? load_global(spos + (d.start + threadIdx.x + B)) : zero4;
  float nw = wscalar;
  if (weight && threadIdx.x < cnt) nw = weight[size_t(__float_as_uint(rec1.w)) * nc + ch];
  for (int i = threadIdx.x; i < kWinCells; i += blockDim.x) win[i] = 0ull;
  const double scale = fixed_point_scale(weight ? *wmax_dev : wmax_host, n);
  bool hit = false;
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < cnt; q += B) {
    const float4 rec = rec1;
    const float wp = nw;
    rec1 = rec2;
    if (weight && q + B < cnt) nw = weight[size_t(__float_as_uint(rec1.w)) * nc + ch];
    if (q + 2 * B < cnt) rec2 = load_global(spos + (d.start + q + 2 * B));
    int lx[2], ly[2], lz[2], gx[2], gy[2], gz[2];
    float wx[2], wy[2], wz[2];
    window_corners<REL>(rec, t, part0, ox, oy, oz, lx, ly, lz, gx, gy, gz, wx, wy, wz);
    if (!REL) {
      // Absolute positions: the 8 corners are the 2^3 block at the floor
      // cell's window index, so each deposit's LDS address is one base plus a
      // compile-time offset.  A particle with a finite weight and position
      // has finite deposits (the CIC weights lie in [0, 1]); the others test
      // each deposit.
      unsigned long long *w0 =
          win + (__umul24(__umul24(uint32_t(lx[0]), uint32_t(kWin)) + uint32_t(ly[0]), uint32_t(kWin)) +
                 uint32_t(lz[0]));
      const bool finite = isfinite(wp) && isfinite(rec.x) && isfinite(rec.y) && isfinite(rec.z);
      if (finite) {
#pragma unroll
        for (int o = 0; o < 8; ++o) {
          const int bx = o & 1, by = (o >> 1) & 1, bz = (o >> 2) & 1;
          const float v = wp * (wx[bx] * wy[by] * wz[bz]);
          atomicAdd(w0 + (bx * kWin + by) * kWin + bz, fixed_point_q(v, scale));
        }
      } else {
#pragma unroll
        for (int o = 0; o < 8; ++o) {
          const int bx = o & 1, by = (o >> 1) & 1, bz = (o >> 2) & 1;
          const float v = wp * (wx[bx] * wy[by] * wz[bz]);
          if (!isfinite(v)) {
            hit = true;
            continue;
          }
          atomicAdd(w0 + (bx * kWin + by) * kWin + bz, fixed_point_q(v, scale));
        }
      }
      continue;
    }
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int bx = o & 1, by = (o >> 1) & 1, bz = (o >> 2) & 1;
      const float wc = wx[bx] * wy[by] * wz[bz];
      const int Lx = lx[bx], Ly = ly[by], Lz = lz[bz];
      if (Lx == -2 || Ly == -2 || Lz == -2) continue;   // dropped corner
      const float v = wp * wc;
      if (Lx >= 0 && Ly >= 0 && Lz >= 0) {
        if (!isfinite(v)) {
          hit = true;
          continue;
        }
        atomicAdd(&win[(Lx * kWin + Ly) * kWin + Lz], fixed_point_q(v, scale));
      } else {
        // rounding put a relative corner outside the window: global path
        const uint32_t cell = (uint32_t(gx[bx]) * uint32_t(t.ny) + uint32_t(gy[by])) * uint32_t(t.nz) + uint32_t(gz[bz]);
        atomicAdd(&overflow[size_t(ch) * M + cell0 + cell], v);
      }
    }
  }
  if (hit) atomicOr(nonfinite, 1u);
  __syncthreads();
  write_window<REL>(win, d, t, itx, ity, itz, scale, (size_t(c) * nc + ch) * kWinCells, partial,
                    mesh + size_t(ch) * M + cell0, accumulate);
}

// --------------------------------------------------------------- reduce

// One workgroup per tile.  Cell (a, b, c) of tile T (local coordinates)
// receives the deposits of T's own windows at (a, b, c) and, on T's low
// faces, the +1 halo of the lower neighbour(s): along an axis where the local
// coordinate is 0, the neighbour at -1 (periodic) contributes at its own
// local index = its tile length.  The (up to 8) sources and each tile's
// chunks are summed in a fixed order.  With DIRECT, a single-chunk tile's
// own deposits are already in the mesh (the accumulation wrote them), and
// only its low-face cells take the neighbours' halos, in the same order.
//
// When the accumulation flagged a non-finite deposit, the workgroup then
// walks the particles of the same (up to 8) source tiles and adds their
// NaN / +-Inf deposits that land in its own tile to the finished cells as
// floats (only this workgroup writes them, so nothing races).
constexpr int kReduceThreads = 256;

template <class G, bool DIRECT>
__global__ void __launch_bounds__(kReduceThreads)
bin_reduce_kernel(const float *__restrict__ partial, const uint32_t *__restrict__ chunk_off,
                  const uint32_t *__restrict__ nchunk, G g, uint32_t M, int nc, const float *__restrict__ overflow,
                  int has_overflow, const float *__restrict__ wmax_dev, float wmax_host, int has_w, uint32_t n,
                  int accumulate, float *__restrict__ mesh, uint32_t K, uint32_t per, bool slabs,
                  const uint32_t *__restrict__ nonfinite, const float4 *__restrict__ spos,
                  const ChunkDesc *__restrict__ desc, const float *__restrict__ weight, float wscalar,
                  FaceCells fc) {
  const uint32_t gtile = slab_item(blockIdx.x, per, slabs);
  if (gtile >= K) return;
  // with fc.fb the finished face cells go there instead of the mesh
  float *fbt = fc.fb ? fc.fb + size_t(blockIdx.y) * fc.chan + size_t(gtile) * kFaceStride : nullptr;
  const unsigned long long *ipart = reinterpret_cast<const unsigned long long *>(partial);
  const double inv = 1.0 / fixed_point_scale(has_w ? *wmax_dev : wmax_host, n);
  const SimRef &sim = g.tile_sim(gtile);
  const TileGrid &t = sim.t;
  const uint32_t tile0 = sim.tile0, cell0 = sim.cell0;
  const uint32_t tile = gtile - tile0;
  const int ch = int(blockIdx.y);
  const uint32_t txy = t.div_tz.div(tile);
  const int itz = int(tile - txy * uint32_t(t.tz));
  const int itx = int(t.div_ty.div(txy));
  const int ity = int(txy - uint32_t(itx) * uint32_t(t.ty));
  const int Lx = tile_len(itx, t.tx, t.nx), Ly = tile_len(ity, t.ty, t.ny), Lz = tile_len(itz, t.tz, t.nz);
  // lower neighbours (periodic) and their lengths
  const int px = itx ? itx - 1 : t.tx - 1, py = ity ? ity - 1 : t.ty - 1, pz = itz ? itz - 1 : t.tz - 1;
  const int PLx = tile_len(px, t.tx, t.nx), PLy = tile_len(py, t.ty, t.ny), PLz = tile_len(pz, t.tz, t.nz);
  // the (up to 8) source tiles, q = sx | sy << 1 | sz << 2 (s = 1: the
  // lower neighbour along that axis): first chunk and chunk count
  __shared__ uint32_t src_off[8], src_cnt[8], src_id[8];
  if (threadIdx.x < 8) {
    const int q = int(threadIdx.x);
    const int jx = (q & 1) ? px : itx, jy = (q & 2) ? py : ity, jz = (q & 4) ? pz : itz;
    const uint32_t src = tile0 + (uint32_t(jx) * uint32_t(t.ty) + uint32_t(jy)) * uint32_t(t.tz) + uint32_t(jz);
    src_off[q] = chunk_off[src];
    src_cnt[q] = nchunk[src];
    src_id[q] = src;
  }
  __syncthreads();
  const bool faces_only = DIRECT && src_cnt[0] == 1;
  const bool full = Lx == kTile && Ly == kTile && Lz == kTile;
  // faces: the a = 0 plane, then b = 0 with a > 0, then c = 0 with a, b > 0
  const int nf0 = Ly * Lz, nf1 = (Lx - 1) * Lz, nf2 = (Lx - 1) * (Ly - 1);
  const int cells_in_tile = faces_only ? nf0 + nf1 + nf2 : Lx * Ly * Lz;
  uint32_t src_max = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) src_max = src_max > src_cnt[q] ? src_max : src_cnt[q];
  if (faces_only && full && PLx == kTile && PLy == kTile && PLz == kTile && src_max <= 1) {
    // The common case: a full single-chunk tile whose lower neighbours are
    // full and hold at most one chunk each.  Each thread loads, for all of
    // its face cells at once, the cell's own value and all 7 neighbour halo
    // entries it may take (those it does not take from the tile's own slot,
    // a valid address), then adds the ones it takes in the order of the
    // general loop below: sources (sx, sy, sz) lexicographically, sz
    // fastest.  Same sums, same bits.
    constexpr int kPer = (kFaces + kReduceThreads - 1) / kReduceThreads;   // face cells per thread
    const float *own = partial + (size_t(src_off[0]) * nc + ch) * kWinCells * 2 + kH;
    float h[kPer][8];
    float *dsts[kPer];
    bool takes[kPer][8];
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int e0 = int(threadIdx.x) + i * kReduceThreads;
      const int e = e0 < kFaces ? e0 : kFaces - 1;   // the last pass repeats a cell it does not store
      int a, b, c;
      face_cell(e, a, b, c);
      const uint32_t cell = cell0 + (uint32_t(itx * kTile + a) * uint32_t(t.ny) + uint32_t(ity * kTile + b)) *
                                        uint32_t(t.nz) + uint32_t(itz * kTile + c);
      dsts[i] = mesh + size_t(ch) * M + cell;
      // the tile's own deposits: from its slot, or, when the paint adds to
      // the mesh, the mesh value the accumulation already added them to
      h[i][0] = accumulate ? *dsts[i] : own[e];
#pragma unroll
      for (int k = 1; k < 8; ++k) {
        const int sx = k >> 2, sy = (k >> 1) & 1, sz = k & 1;
        const int q = sx | (sy << 1) | (sz << 2);
        takes[i][k] = (!sx || a == 0) && (!sy || b == 0) && (!sz || c == 0) && src_cnt[q] == 1;
        const uint32_t slot = takes[i][k] ? src_off[q] : src_off[0];
        h[i][k] = partial[((size_t(slot) * nc + ch) * kWinCells) * 2 +
                          halo_index(sx ? kTile : a, sy ? kTile : b, sz ? kTile : c)];
      }
    }
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      if (int(threadIdx.x) + i * kReduceThreads >= kFaces) continue;
      float acc = h[i][0];
#pragma unroll
      for (int k = 1; k < 8; ++k) acc = takes[i][k] ? acc + h[i][k] : acc;
      if (fbt)
        fbt[int(threadIdx.x) + i * kReduceThreads] = acc;
      else
        *dsts[i] = acc;
    }
  } else {
    for (int e = threadIdx.x; e < cells_in_tile; e += blockDim.x) {
      int a, b, c;
      if (full) {
        // the same enumeration with the divisors known at compile time
        if (!faces_only) {
          a = e >> 8; b = (e >> 4) & 15; c = e & 15;
        } else {
          face_cell(e, a, b, c);
        }
      } else if (!faces_only) {
        c = e % Lz;
        const int ab = e / Lz;
        b = ab % Ly;
        a = ab / Ly;
      } else if (e < nf0) {
        a = 0; b = e / Lz; c = e % Lz;
      } else if (e < nf0 + nf1) {
        const int f = e - nf0;
        a = 1 + f / Lz; b = 0; c = f % Lz;
      } else {
        const int f = e - nf0 - nf1;
        a = 1 + f / (Ly - 1); b = 1 + f % (Ly - 1); c = 0;
      }
      const uint32_t cell = cell0 + (uint32_t(itx * kTile + a) * uint32_t(t.ny) + uint32_t(ity * kTile + b)) *
                                        uint32_t(t.nz) + uint32_t(itz * kTile + c);
      float *dst = mesh + size_t(ch) * M + cell;
      // a full single-chunk tile keeps its own face cells in its slot too
      float acc = 0.0f;
      if (faces_only)
        acc = full && !accumulate ? partial[(size_t(src_off[0]) * nc + ch) * kWinCells * 2 + kH + e] : *dst;
      const int nxs = a == 0 ? 2 : 1, nys = b == 0 ? 2 : 1, nzs = c == 0 ? 2 : 1;
      for (int sx = 0; sx < nxs; ++sx)
        for (int sy = 0; sy < nys; ++sy)
          for (int sz = 0; sz < nzs; ++sz) {
            if (faces_only && !(sx | sy | sz)) continue;  // own deposits: already in acc
            const int q = sx | (sy << 1) | (sz << 2);
            const int la = sx ? PLx : a, lb = sy ? PLy : b, lc = sz ? PLz : c;
            const int li = (la * kWin + lb) * kWin + lc;
            const uint32_t c0 = src_off[q], cn = src_cnt[q];
            if (cn == 1) {
              // a full single-chunk source (absolute mode) holds its halo compactly
              const bool compact = DIRECT && (sx ? PLx : Lx) == kTile && (sy ? PLy : Ly) == kTile &&
                                   (sz ? PLz : Lz) == kTile;
              acc += partial[((size_t(c0) * nc + ch) * kWinCells) * 2 + (compact ? halo_index(la, lb, lc) : li)];
            } else if (cn > 1) {
              unsigned long long sum = 0ull;
              for (uint32_t k = 0; k < cn; ++k) sum += ipart[(size_t(c0 + k) * nc + ch) * kWinCells + li];
              acc += float(double((long long)sum) * inv);
            }
          }
      if (faces_only) {
        if (fbt)
          fbt[e] = acc;   // a full tile: e is the face index
        else
          *dst = acc;
        continue;
      }
      float v = acc;
      if (has_overflow) v += overflow[size_t(ch) * M + cell];
      if (fbt && (a == 0 || b == 0 || c == 0))
        fbt[face_index(a, b, c)] = v;
      else
        *dst = accumulate ? *dst + v : v;
    }
  }
  if (*nonfinite == 0u) return;
  __threadfence();
  __syncthreads();
  float *meshc = mesh + size_t(ch) * M + cell0;
  for (int q = 0; q < 8; ++q) {
    bool dup = false;
    for (int r = 0; r < q; ++r) dup |= src_id[r] == src_id[q];
    if (dup) continue;   // a tile grid one or two tiles long along an axis
    for (uint32_t k = 0; k < src_cnt[q]; ++k) {
      const ChunkDesc dk = desc[src_off[q] + k];
      int sx, sy, sz;
      tile_origin(dk.tile - tile0, t, sx, sy, sz);
      const int ox = sx * kTile, oy = sy * kTile, oz = sz * kTile;
      for (uint32_t i = threadIdx.x; i < dk.count; i += blockDim.x) {
        const float4 rec = spos[dk.start + i];
        const float wp = weight ? weight[size_t(__float_as_uint(rec.w)) * nc + ch] : wscalar;
        int lx[2], ly[2], lz[2], gx[2], gy[2], gz[2];
        float wx[2], wy[2], wz[2];
        window_corners<!DIRECT>(rec, t, sim.part0, ox, oy, oz, lx, ly, lz, gx, gy, gz, wx, wy, wz);
        for (int o = 0; o < 8; ++o) {
          const int bx = o & 1, by = (o >> 1) & 1, bz = (o >> 2) & 1;
          const float v = wp * (wx[bx] * wy[by] * wz[bz]);
          if (isfinite(v)) continue;
          int cx, cy, cz;
          if (DIRECT) {
            cx = ox + lx[bx];
            cy = oy + ly[by];
            cz = oz + lz[bz];
            cx -= cx >= t.nx ? t.nx : 0;
            cy -= cy >= t.ny ? t.ny : 0;
            cz -= cz >= t.nz ? t.nz : 0;
          } else {
            // dropped corners, and those the overflow mesh carried
            if (lx[bx] < 0 || ly[by] < 0 || lz[bz] < 0) continue;
            cx = gx[bx];
            cy = gy[by];
            cz = gz[bz];
          }
          if (cx / kTile != itx || cy / kTile != ity || cz / kTile != itz) continue;
          const int fa = cx - itx * kTile, fbb = cy - ity * kTile, fcc = cz - itz * kTile;
          if (fbt && (fa == 0 || fbb == 0 || fcc == 0))
            atomicAdd(&fbt[face_index(fa, fbb, fcc)], v);
          else
            atomicAdd(&meshc[(uint32_t(cx) * uint32_t(t.ny) + uint32_t(cy)) * uint32_t(t.nz) + uint32_t(cz)], v);
        }
      }
    }
  }
}

}  // namespace jpm

namespace jpm {

int abs_max(hipStream_t s, const float *w, uint32_t n, float *out);
int zero_f32(hipStream_t s, float *p, int64_t m);

template <typename K>
static int set_lds(K kernel, size_t bytes) {
  if (bytes > 65536)
    JPM_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                int(bytes)));
  return JPM_OK;
}

// Sizes of a tile sort (batch totals): tiles, particles, count/scatter
// workgroups of `slice` particles, chunk-table capacity, and the largest
// per-simulation tile count (the LDS histogram size).
struct SortDims {
  uint32_t K, n, G, slice, chunk, maxchunks, maxK, B;
  bool slabs;
};

static SortDims one_dims(const TileGrid &t, int64_t n, uint32_t flags) {
  SortDims d;
  const uint64_t nn = uint64_t(n > 0 ? n : 1);
  d.K = t.K;
  d.n = uint32_t(n);
  d.slice = tuned_slice(flags);
  d.chunk = tuned_chunk(flags);
  d.G = uint32_t((nn + d.slice - 1) / d.slice);
  d.maxchunks = uint32_t((nn + d.chunk - 1) / d.chunk + t.K);
  d.maxK = t.K;
  d.B = 1;
  d.slabs = tuned_slabs(flags);
  return d;
}

// Workspace of the tile sort (byte offsets).
struct SortWs {
  size_t count, tstart, cursor, nchunk, choff, sums, totals, desc, spos, total;
};

static SortWs sort_ws(const SortDims &d) {
  SortWs w;
  const size_t nn = size_t(d.n > 0 ? d.n : 1);
  const size_t nsums = (size_t(d.K) + kScanBlock - 1) / kScanBlock + 1;
  size_t o = 0;
  w.count = o;    o += align_up(size_t(d.K) * 4);
  w.tstart = o;   o += align_up(size_t(d.K) * 4);
  w.cursor = o;   o += align_up(size_t(d.K) * 4);
  w.nchunk = o;   o += align_up(size_t(d.K) * 4);
  w.choff = o;    o += align_up(size_t(d.K) * 4);
  w.sums = o;     o += align_up(nsums * 4);
  w.totals = o;   o += align_up(16);
  w.desc = o;     o += align_up(size_t(d.maxchunks) * sizeof(ChunkDesc));
  w.spos = o;     o += align_up(nn * 16);
  w.total = o;
  return w;
}

// Workspace of the window accumulation that follows a sort.
static size_t accum_ws(const SortDims &d, size_t M, int nc, bool rel, size_t *partial, size_t *overflow,
                       size_t *wmax) {
  size_t o = 0;
  *partial = o;  o += align_up(size_t(d.maxchunks) * nc * kWinCells * 8);
  *overflow = o; o += rel ? align_up(size_t(nc) * M * 4) : 0;
  *wmax = o;     o += align_up(4);
  return o;
}

bool binned_paint_supported(int nx, int ny, int nz) {
  return nx > 0 && ny > 0 && nz > 0 && 3 * cells(nx, ny, nz) <= (int64_t(1) << 32);
}

size_t bin_sort_workspace_size(int nx, int ny, int nz, int64_t n, uint32_t flags) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
  return sort_ws(one_dims(make_tiles(nx, ny, nz), n, flags)).total;
}

size_t bins_paint_workspace_size(int nx, int ny, int nz, int64_t n, int nc, uint32_t flags) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || nc < 1) return 0;
  size_t a, b, c;
  return accum_ws(one_dims(make_tiles(nx, ny, nz), n, flags), size_t(cells(nx, ny, nz)), nc, flags & JPM_RELATIVE, &a, &b,
                  &c);
}

size_t binned_paint_workspace_size(int nx, int ny, int nz, int64_t n, int nc, uint32_t flags) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || nc < 1) return 0;
  return align_up(bin_sort_workspace_size(nx, ny, nz, n, flags)) + bins_paint_workspace_size(nx, ny, nz, n, nc, flags);
}

template <class G>
static BinsT<G> view_of(const G &g, const SortDims &d, uint32_t M, bool rel, const void *ws) {
  const SortWs w = sort_ws(d);
  const char *base = static_cast<const char *>(ws);
  BinsT<G> b;
  b.g = g;
  b.K = d.K;
  b.n = d.n;
  b.M = M;
  b.maxchunks = d.maxchunks;
  b.chunk = d.chunk;
  b.rel = rel;
  b.slabs = d.slabs;
  b.spos = reinterpret_cast<const float4 *>(base + w.spos);
  b.desc = reinterpret_cast<const ChunkDesc *>(base + w.desc);
  b.nchunk = reinterpret_cast<const uint32_t *>(base + w.nchunk);
  b.choff = reinterpret_cast<const uint32_t *>(base + w.choff);
  b.total_chunks = reinterpret_cast<const uint32_t *>(base + w.totals) + 1;
  b.nonfinite = const_cast<uint32_t *>(reinterpret_cast<const uint32_t *>(base + w.totals) + 2);
  return b;
}

// The Bins view of a sort already held in `ws` (same shape, count, flags).
Bins bins_view(int nx, int ny, int nz, int64_t n, uint32_t flags, const void *ws) {
  const TileGrid t = make_tiles(nx, ny, nz);
  OneSim g{one_sim(nx, ny, nz, n)};
  return view_of(g, one_dims(t, n, flags), uint32_t(cells(nx, ny, nz)), flags & JPM_RELATIVE, ws);
}

// Sort the particles of every simulation into its tiles (workspace of
// sort_ws(d).total bytes).  counts_in: NULL, or the tile counts of these
// positions already (d.K entries, e.g. from the kick that drifted them),
// which skips the counting pass.
template <class G>
static int sort_into(hipStream_t s, const G &g, const SortDims &d, uint32_t M, bool rel, const float *pos, void *ws,
                     size_t ws_bytes, BinsT<G> &out, const uint32_t *counts_in = nullptr) {
  const SortWs w = sort_ws(d);
  JPM_REQUIRE(ws && ws_bytes >= w.total, "tile sort needs %zu workspace bytes, got %zu", w.total, ws_bytes);
  JPM_REQUIRE(pos != nullptr, "positions are null");
  const bool lds = d.maxK <= uint32_t(kMaxKeysLds);   // per-workgroup LDS histograms fit
  char *base = static_cast<char *>(ws);
  uint32_t *count = reinterpret_cast<uint32_t *>(base + w.count);
  uint32_t *tstart = reinterpret_cast<uint32_t *>(base + w.tstart);
  uint32_t *cursor = reinterpret_cast<uint32_t *>(base + w.cursor);
  uint32_t *nchunk = reinterpret_cast<uint32_t *>(base + w.nchunk);
  uint32_t *choff = reinterpret_cast<uint32_t *>(base + w.choff);
  uint32_t *sums = reinterpret_cast<uint32_t *>(base + w.sums);
  uint32_t *totals = reinterpret_cast<uint32_t *>(base + w.totals);  // [0] particles, [1] chunks
  const size_t lds_hist = size_t(d.maxK) * 4;
  int rc;

  // 1. tile counts
  if (counts_in) {
    count = const_cast<uint32_t *>(counts_in);   // only read below
  } else {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(grid_for(d.K, 256)), dim3(256), 0, s, count, d.K);
    JPM_LAUNCHED("zero_u32_kernel");
    if (lds) {
      auto kern = rel ? bin_count_kernel<true, G> : bin_count_kernel<false, G>;
      rc = set_lds(kern, lds_hist);
      if (rc) return rc;
      JPM_TIMED(s, "bin_count_kernel");
      hipLaunchKernelGGL(kern, dim3(d.G), dim3(kBinThreads), lds_hist, s, pos, d.slice, g, count);
      JPM_LAUNCHED("bin_count_kernel");
    } else {
      auto kern = rel ? bin_count_global_kernel<true, G> : bin_count_global_kernel<false, G>;
      JPM_TIMED(s, "bin_count_kernel");
      hipLaunchKernelGGL(kern, dim3(d.G), dim3(kBinThreads), 0, s, pos, d.slice, g, count);
      JPM_LAUNCHED("bin_count_global_kernel");
    }
  }
  // 2. tile starts, scatter cursors, chunk table
  if (d.K <= kTableMaxK) {
    hipLaunchKernelGGL(tile_table_kernel, dim3((d.K + kTableThreads - 1) / kTableThreads), dim3(kTableThreads), 0, s,
                       count, d.K, d.chunk, tstart, cursor,
                       nchunk, choff, totals, reinterpret_cast<ChunkDesc *>(base + w.desc));
    JPM_LAUNCHED("tile_table_kernel");
  } else {
    JPM_HIP(hipMemcpyAsync(tstart, count, size_t(d.K) * 4, hipMemcpyDeviceToDevice, s));
    rc = exclusive_scan(s, tstart, d.K, sums, totals);
    if (rc) return rc;
    JPM_HIP(hipMemcpyAsync(cursor, tstart, size_t(d.K) * 4, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(tile_chunks_kernel, dim3(grid_for(d.K, 256)), dim3(256), 0, s, count, d.K, d.chunk, nchunk);
    JPM_LAUNCHED("tile_chunks_kernel");
    JPM_HIP(hipMemcpyAsync(choff, nchunk, size_t(d.K) * 4, hipMemcpyDeviceToDevice, s));
    rc = exclusive_scan(s, choff, d.K, sums, totals + 1);
    if (rc) return rc;
    hipLaunchKernelGGL(chunk_desc_kernel, dim3(grid_for(d.K, 256)), dim3(256), 0, s, tstart, count, choff, d.K,
                       d.chunk, reinterpret_cast<ChunkDesc *>(base + w.desc));
    JPM_LAUNCHED("chunk_desc_kernel");
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, totals + 2, 1u);
    JPM_LAUNCHED("zero_u32_kernel");
  }
  // 3. scatter particles into tile order
  if (lds && (d.slice == 4u * kScatterRegThreads || d.slice == 8u * kScatterRegThreads ||
              d.slice == 16u * kScatterRegThreads)) {
    const int v = d.slice == 4u * kScatterRegThreads ? 0 : d.slice == 8u * kScatterRegThreads ? 1 : 2;
    auto kern = rel ? (v == 0 ? bin_scatter_reg_kernel<true, G, 8, 512>
                              : v == 1 ? bin_scatter_reg_kernel<true, G, 8, 1024> : bin_scatter_reg_kernel<true, G, 16, 1024>)
                    : (v == 0 ? bin_scatter_reg_kernel<false, G, 8, 512>
                              : v == 1 ? bin_scatter_reg_kernel<false, G, 8, 1024> : bin_scatter_reg_kernel<false, G, 16, 1024>);
    rc = set_lds(kern, lds_hist);
    if (rc) return rc;
    JPM_TIMED(s, "bin_scatter_kernel");
    hipLaunchKernelGGL(kern, dim3(d.G), dim3(v == 0 ? 512 : kScatterRegThreads), lds_hist, s, pos, d.slice, g,
                       cursor, reinterpret_cast<float4 *>(base + w.spos));
    JPM_LAUNCHED("bin_scatter_reg_kernel");
  } else if (lds) {
    auto kern = rel ? bin_scatter_kernel<true, G> : bin_scatter_kernel<false, G>;
    rc = set_lds(kern, lds_hist);
    if (rc) return rc;
    JPM_TIMED(s, "bin_scatter_kernel");
    hipLaunchKernelGGL(kern, dim3(d.G), dim3(kBinThreads), lds_hist, s, pos, d.slice, g, cursor,
                       reinterpret_cast<float4 *>(base + w.spos));
    JPM_LAUNCHED("bin_scatter_kernel");
  } else {
    auto kern = rel ? bin_scatter_global_kernel<true, G> : bin_scatter_global_kernel<false, G>;
    JPM_TIMED(s, "bin_scatter_kernel");
    hipLaunchKernelGGL(kern, dim3(d.G), dim3(kBinThreads), 0, s, pos, d.slice, g, cursor,
                       reinterpret_cast<float4 *>(base + w.spos));
    JPM_LAUNCHED("bin_scatter_global_kernel");
  }
  out = view_of(g, d, M, rel, ws);
  return JPM_OK;
}

// Sort n particles into tiles (workspace `ws` of bin_sort_workspace_size).
int bin_sort(hipStream_t s, int nx, int ny, int nz, const float *pos, int64_t n, uint32_t flags, void *ws,
             size_t ws_bytes, Bins &out, const uint32_t *counts_in) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(n >= 1 && n <= kMaxCells, "particle count %lld out of range", (long long)n);
  const TileGrid t = make_tiles(nx, ny, nz);
  const int64_t M = cells(nx, ny, nz);
  JPM_REQUIRE(3 * M <= (int64_t(1) << 32), "mesh too large");
  const bool rel = flags & JPM_RELATIVE;
  JPM_REQUIRE(!rel || n == M, "relative mode needs one particle per lattice node (n=%lld, mesh %lld)",
              (long long)n, (long long)M);
  OneSim g{one_sim(nx, ny, nz, n)};
  return sort_into(s, g, one_dims(t, n, flags), uint32_t(M), rel, pos, ws, ws_bytes, out, counts_in);
}

// order[k] = index of the particle the tile sort put at slot k: the index
// bits of the sorted records.
__global__ void __launch_bounds__(256)
record_order_kernel(const float4 *__restrict__ spos, uint32_t n, uint32_t *__restrict__ order) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) order[k] = __float_as_uint(spos[k].w);
}

template <class G>
int bins_order(hipStream_t s, const BinsT<G> &b, uint32_t *order) {
  JPM_REQUIRE(order, "null pointer argument");
  JPM_TIMED(s, "record_order_kernel");
  hipLaunchKernelGGL(record_order_kernel, dim3((b.n + 255) / 256), dim3(256), 0, s, b.spos, b.n, order);
  JPM_LAUNCHED("record_order_kernel");
  return JPM_OK;
}
template int bins_order<OneSim>(hipStream_t, const Bins &, uint32_t *);
template int bins_order<SimTable>(hipStream_t, const BatchBins &, uint32_t *);

// nc meshes (stride b.M) = [mesh +] window sums of the sorted particles,
// channel c with weight[p*nc + c] (or wscalar when weight is NULL).
// wmax_in: NULL, or a device float holding max |weight| already (the
// fixed-point scale's bound), which skips that pass.
template <class G>
static int paint_sorted(hipStream_t s, const BinsT<G> &b, const float *weight, float wscalar, int nc, float *mesh,
                        uint32_t flags, void *ws, size_t ws_bytes, const float *wmax_in = nullptr,
                        const FaceCells *faces = nullptr) {
  SortDims d{};
  d.maxchunks = b.maxchunks;
  JPM_REQUIRE(nc == 1 || nc == 3, "binned paint supports 1 or 3 channels (got %d)", nc);
  JPM_REQUIRE(mesh != nullptr, "mesh is null");
  const int64_t M = b.M;
  JPM_REQUIRE(int64_t(nc) * M <= (int64_t(1) << 32), "mesh too large");
  size_t opart, oover, owmax;
  const size_t need = accum_ws(d, size_t(M), nc, b.rel, &opart, &oover, &owmax);
  JPM_REQUIRE(ws && ws_bytes >= need, "binned paint needs %zu workspace bytes, got %zu", need, ws_bytes);
  const bool accumulate = !(flags & JPM_ZERO_MESH);
  const FaceCells fc = faces ? *faces : FaceCells{nullptr, 0, 0, 0, 0, FastDiv(1)};
  JPM_REQUIRE(!fc.fb || (!b.rel && !accumulate && size_t(b.K) * kFaceStride <= fc.chan),
              "face cells apart need absolute positions, a replaced mesh and room for every tile");
  char *base = static_cast<char *>(ws);
  int rc;
  if (b.rel) {
    rc = zero_f32(s, reinterpret_cast<float *>(base + oover), nc * M);
    if (rc) return rc;
  }
  const float wmax_host = finite_abs(wscalar);
  if (weight && !wmax_in) {
    rc = abs_max(s, weight, uint32_t(int64_t(b.n) * nc), reinterpret_cast<float *>(base + owmax));
    if (rc) return rc;
  }
  auto *partial = reinterpret_cast<float *>(base + opart);
  const float *wmax_dev = wmax_in ? wmax_in : reinterpret_cast<const float *>(base + owmax);
  float *overflow = reinterpret_cast<float *>(base + oover);
  {
    JPM_TIMED(s, nc == 1 ? "bin_accum_kernel[nc=1]" : "bin_accum_kernel[nc=3]");
    auto kern = b.rel ? bin_accum_kernel<true, G> : bin_accum_kernel<false, G>;
    uint32_t per;
    (void)slab_grid(b.maxchunks, b.slabs, per);
    const unsigned grid = 8u * per * unsigned(nc);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kWindowThreads), 0, s, b.spos, weight, nc, wscalar, b.desc,
                       b.total_chunks, b.g, b.M, wmax_dev, wmax_host, b.n, partial, overflow, int(accumulate), mesh,
                       per, b.slabs, b.nonfinite);
  }
  JPM_LAUNCHED("bin_accum_kernel");
  {
    JPM_TIMED(s, nc == 1 ? "bin_reduce_kernel[nc=1]" : "bin_reduce_kernel[nc=3]");
    // absolute positions: single-chunk tiles wrote their own cells already
    auto kern = b.rel ? bin_reduce_kernel<G, false> : bin_reduce_kernel<G, true>;
    // always in XCD slabs: tiles consecutive in z then run on one XCD, and
    // their low-z faces, single floats a tile apart along each mesh row,
    // share that XCD's L2 lines (0.14 -> 0.10 ms for 3 channels at 256^3)
    uint32_t per;
    const unsigned grid = slab_grid(b.K, true, per);
    hipLaunchKernelGGL(kern, dim3(grid, nc), dim3(kReduceThreads), 0, s, partial, b.choff, b.nchunk, b.g, b.M, nc, overflow,
                       int(b.rel), wmax_dev, wmax_host, int(weight != nullptr), b.n, int(accumulate), mesh, b.K, per,
                       true, b.nonfinite, b.spos, b.desc, weight, wscalar, fc);
  }
  JPM_LAUNCHED("bin_reduce_kernel");
  return JPM_OK;
}

int bins_paint(hipStream_t s, const Bins &b, const float *weight, float wscalar, int nc, float *mesh, uint32_t flags,
               void *ws, size_t ws_bytes, const float *wmax_in, const FaceCells *faces) {
  return paint_sorted(s, b, weight, wscalar, nc, mesh, flags, ws, ws_bytes, wmax_in, faces);
}

int binned_paint(hipStream_t s, int nx, int ny, int nz, const float *pos, int64_t n, const float *weight,
                 float wscalar, int nc, float *mesh, uint32_t flags, void *ws, size_t ws_bytes) {
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  JPM_REQUIRE(nc == 1 || nc == 3, "binned paint supports 1 or 3 channels (got %d)", nc);
  JPM_REQUIRE(n >= 0 && n <= kMaxCells, "particle count %lld out of range", (long long)n);
  JPM_REQUIRE(mesh != nullptr, "mesh is null");
  const int64_t M = cells(nx, ny, nz);
  JPM_REQUIRE(int64_t(nc) * M <= (int64_t(1) << 32), "mesh too large");
  const size_t need = binned_paint_workspace_size(nx, ny, nz, n, nc, flags);
  JPM_REQUIRE(ws && ws_bytes >= need, "binned paint needs %zu workspace bytes, got %zu", need, ws_bytes);
  if (n == 0) {
    if (flags & JPM_ZERO_MESH) return zero_f32(s, mesh, nc * M);
    return JPM_OK;
  }
  const size_t sort_bytes = align_up(bin_sort_workspace_size(nx, ny, nz, n, flags));
  Bins b;
  rc = bin_sort(s, nx, ny, nz, pos, n, flags, ws, sort_bytes, b);
  if (rc) return rc;
  return bins_paint(s, b, weight, wscalar, nc, mesh, flags, static_cast<char *>(ws) + sort_bytes, ws_bytes - sort_bytes,
                    nullptr);
}

// ------------------------------------------------------------- batches

static SortDims batch_dims(const SimBatch &bt) {
  SortDims d;
  d.K = bt.K;
  d.n = bt.N;
  d.slice = bt.slice;
  d.chunk = bt.chunk;
  d.G = bt.G;
  d.maxchunks = uint32_t((uint64_t(bt.N) + bt.chunk - 1) / bt.chunk + bt.K + uint32_t(bt.B));
  d.maxK = bt.maxK;
  d.B = uint32_t(bt.B);
  d.slabs = bt.slabs;
  return d;
}

size_t batch_sort_workspace_size(const SimBatch &bt) { return sort_ws(batch_dims(bt)).total; }

size_t batch_paint_workspace_size(const SimBatch &bt, int nc) {
  size_t a, b, c;
  return accum_ws(batch_dims(bt), bt.M, nc, bt.rel, &a, &b, &c);
}

int batch_sort(hipStream_t s, const SimBatch &bt, const float *pos, void *ws, size_t ws_bytes, BatchBins &out,
               const uint32_t *counts_in) {
  SimTable g{bt.dev, bt.B};
  return sort_into(s, g, batch_dims(bt), bt.M, bt.rel, pos, ws, ws_bytes, out, counts_in);
}

int zero_u32(hipStream_t s, uint32_t *p, uint32_t n) {
  hipLaunchKernelGGL(zero_u32_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n);
  JPM_LAUNCHED("zero_u32_kernel");
  return JPM_OK;
}

BatchBins batch_view(const SimBatch &bt, const void *ws) {
  SimTable g{bt.dev, bt.B};
  return view_of(g, batch_dims(bt), bt.M, bt.rel, ws);
}

int batch_paint(hipStream_t s, const SimBatch &bt, const BatchBins &b, const float *weight, int nc, float *mesh,
                void *ws, size_t ws_bytes, const float *wmax_in) {
  return paint_sorted(s, b, weight, 1.0f, nc, mesh, JPM_ZERO_MESH, ws, ws_bytes, wmax_in);
}

}  // namespace jpm
