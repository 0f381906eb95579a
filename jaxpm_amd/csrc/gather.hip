// Window gathers over the tile sort: the force read (fused with the kick or
// writing the forces) and the position cotangent of the force, both taking
// the *potential* phi and applying the 4-point gradient stencil in LDS.
//
// One workgroup per chunk of a tile's particles.  The workgroup loads the
// potential over the tile plus a 2-cell stencil halo, [o-2, o+19)^3 (21^3
// floats, periodic), into LDS and turns it into the force window: the
// 4-point gradient -D_d phi(c) = -(8 (phi(c+1) - phi(c-1)) - (phi(c+2) -
// phi(c-2))) / 12 at every cell c of [o, o+17)^3, where each particle's 8
// CIC corners lie.  Results go back to the particle's original index.
// Relative-mode corners that rounding puts outside the window (see
// lcorner_rel) read the mesh directly.
#include "tiles.hpp"

namespace jpm {

constexpr int kHalo = 2;
constexpr int kPw = kWin + 2 * kHalo;        // 21
constexpr int kPwCells = kPw * kPw * kPw;    // 9261

struct CornerSet {
  int l[3][2];      // window-local index per axis and corner bit (-1 outside, -2 dropped)
  int g[3][2];      // global cell per axis and corner bit (relative mode fallback)
  float w[3][2];    // CIC weight factors
  float d[3][2];    // their derivatives
};

template <bool REL>
__device__ __forceinline__ void corner_set(float4 rec, const TileGrid &t, uint32_t part0, int ox, int oy, int oz,
                                           CornerSet &cs) {
  const float x = rec.x, y = rec.y, z = rec.z;
  if (REL) {
    const uint32_t pidx = __float_as_uint(rec.w) - part0;
    uint32_t r = t.div_nz.div(pidx);
    uint32_t k = pidx - r * uint32_t(t.nz);
    uint32_t i = t.div_ny.div(r);
    uint32_t j = r - i * uint32_t(t.ny);
    const float p[3] = {float(i) + x, float(j) + y, float(k) + z};
    const int n[3] = {t.nx, t.ny, t.nz};
    const int o[3] = {ox, oy, oz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        lcorner_rel(p[a], float(b), n[a], o[a], cs.l[a][b], cs.w[a][b], cs.g[a][b]);
        // the same nd the weight came from gives the derivative sign
        const float L = float(n[a]);
        float q = floorf(jmod(p[a] + float(b), L));
        float nd = p[a] - q;
        nd = nd - rintf(nd / L) * L;
        cs.d[a][b] = cs.l[a][b] == -2 ? 0.0f : (nd >= 0.0f ? -1.0f : 1.0f);
      }
    }
  } else {
    const float p[3] = {x, y, z};
    const int n[3] = {t.nx, t.ny, t.nz};
    const int o[3] = {ox, oy, oz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      LAxis ax = laxis_abs(p[a], o[a], n[a]);
      cs.l[a][0] = ax.l0;
      cs.l[a][1] = ax.l1;
      cs.w[a][0] = ax.w0;
      cs.w[a][1] = ax.w1;
      const float f = floorf(p[a]);
      cs.d[a][0] = (p[a] - f) >= 0.0f ? -1.0f : 1.0f;
      cs.d[a][1] = (p[a] - (f + 1.0f)) >= 0.0f ? -1.0f : 1.0f;
      cs.g[a][0] = cs.g[a][1] = 0;
    }
  }
}

__device__ __forceinline__ int wrapn(int c, int n) {
  if (uint32_t(c) < uint32_t(n)) return c;
  int r = c % n;
  return r < 0 ? r + n : r;
}

// c mod n for c in [-n, 2n): one compare-and-add instead of a division.
__device__ __forceinline__ int wrap_near(int c, int n) { return c < 0 ? c + n : (c >= n ? c - n : c); }

// Every window coordinate o - halo + a (o < n, halo <= 2, a < 21) lies in
// [-n, 2n) once each mesh edge is at least 21 cells.
__device__ __forceinline__ bool near_wrap_ok(const TileGrid &t) { return t.nx >= kPw && t.ny >= kPw && t.nz >= kPw; }

// Load phi over [o - HALO, o - HALO + EDGE)^3 (periodic) into win, with
// every load of a thread in flight together: the thread computes the
// addresses of all its cells, loads them into registers, and only then
// writes them to LDS.  (Writing each cell as it arrives made the compiler
// wait on every load before issuing the next one.)  The launch has T
// threads.  With edges of at least kPw cells the periodic wrap
// is one compare-and-add per axis, without branches.
template <int EDGE, int HALO, bool NEAR, int T = kWindowThreads>
__device__ __forceinline__ void load_window_batched(const float *__restrict__ phi, const TileGrid &t, int ox, int oy,
                                                    int oz, float *win) {
  constexpr int CELLS = EDGE * EDGE * EDGE;
  constexpr int PER = (CELLS + T - 1) / T;
  const uint32_t ny = uint32_t(t.ny), nz = uint32_t(t.nz);
  float v[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e0 = int(threadIdx.x) + i * T;
    const int e = e0 < CELLS ? e0 : CELLS - 1;   // the last pass repeats a cell it does not store
    const int c = e % EDGE, ab = e / EDGE;
    const int b = ab % EDGE, a = ab / EDGE;
    int gx = ox - HALO + a, gy = oy - HALO + b, gz = oz - HALO + c;
    if (NEAR) {
      gx = wrap_near(gx, t.nx);
      gy = wrap_near(gy, t.ny);
      gz = wrap_near(gz, t.nz);
    } else {
      gx = wrapn(gx, t.nx);
      gy = wrapn(gy, t.ny);
      gz = wrapn(gz, t.nz);
    }
    v[i] = phi[(uint32_t(gx) * ny + uint32_t(gy)) * nz + uint32_t(gz)];
  }
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e0 = int(threadIdx.x) + i * T;
    if (e0 < CELLS) win[e0] = v[i];
  }
}

// Two meshes' windows with all loads of a thread in flight together: a
// over [o - HA, o - HA + EA)^3 into wa, b over [o - HB, o - HB + EB)^3
// into wb, the same cells and order as two load_window_batched calls.
template <int EA, int HA, int EB, int HB, bool NEAR, int T = kWindowThreads>
__device__ __forceinline__ void load_window_pair(const float *__restrict__ a, float *wa, const float *__restrict__ b,
                                                 float *wb, const TileGrid &t, int ox, int oy, int oz) {
  constexpr int CA = EA * EA * EA, CB = EB * EB * EB;
  constexpr int PA = (CA + T - 1) / T, PB = (CB + T - 1) / T;
  const uint32_t ny = uint32_t(t.ny), nz = uint32_t(t.nz);
  auto cell = [&](int e, int EDGE, int HALO) {
    const int c = e % EDGE, ab = e / EDGE;
    const int bb = ab % EDGE, aa = ab / EDGE;
    int gx = ox - HALO + aa, gy = oy - HALO + bb, gz = oz - HALO + c;
    if (NEAR) {
      gx = wrap_near(gx, t.nx);
      gy = wrap_near(gy, t.ny);
      gz = wrap_near(gz, t.nz);
    } else {
      gx = wrapn(gx, t.nx);
      gy = wrapn(gy, t.ny);
      gz = wrapn(gz, t.nz);
    }
    return (uint32_t(gx) * ny + uint32_t(gy)) * nz + uint32_t(gz);
  };
  float va[PA], vb[PB];
#pragma unroll
  for (int i = 0; i < PA; ++i) {
    const int e0 = int(threadIdx.x) + i * T;
    va[i] = a[cell(e0 < CA ? e0 : CA - 1, EA, HA)];
  }
#pragma unroll
  for (int i = 0; i < PB; ++i) {
    const int e0 = int(threadIdx.x) + i * T;
    vb[i] = b[cell(e0 < CB ? e0 : CB - 1, EB, HB)];
  }
#pragma unroll
  for (int i = 0; i < PA; ++i) {
    const int e0 = int(threadIdx.x) + i * T;
    if (e0 < CA) wa[e0] = va[i];
  }
#pragma unroll
  for (int i = 0; i < PB; ++i) {
    const int e0 = int(threadIdx.x) + i * T;
    if (e0 < CB) wb[e0] = vb[i];
  }
}

template <int EDGE, int HALO, int T = kWindowThreads>
__device__ __forceinline__ void load_window_t(const float *__restrict__ phi, const TileGrid &t, int ox, int oy, int oz,
                                              float *win) {
  if (near_wrap_ok(t))
    load_window_batched<EDGE, HALO, true, T>(phi, t, ox, oy, oz, win);
  else
    load_window_batched<EDGE, HALO, false, T>(phi, t, ox, oy, oz, win);
}

constexpr float kC8 = 8.0f / 12.0f;
constexpr float kC1 = 1.0f / 12.0f;

// -D phi at window cell (a, b, c) (already offset by the halo), all 3 axes.
__device__ __forceinline__ void stencil_lds(const float *win, int a, int b, int c, float F[3]) {
  const int sx = kPw * kPw, sy = kPw;
  const int e = (a * kPw + b) * kPw + c;
  F[0] = -(kC8 * (win[e + sx] - win[e - sx]) - kC1 * (win[e + 2 * sx] - win[e - 2 * sx]));
  F[1] = -(kC8 * (win[e + sy] - win[e - sy]) - kC1 * (win[e + 2 * sy] - win[e - 2 * sy]));
  F[2] = -(kC8 * (win[e + 1] - win[e - 1]) - kC1 * (win[e + 2] - win[e - 2]));
}

// The force window: -D_d phi at every cell of the tile's 17^3 window,
// fw[d * kWinCells + (a * kWin + b) * kWin + c], from the 21^3 potential
// window held at the start of the same LDS buffer (3 * kWinCells floats).
// Each force is stencil_lds's expression, so a particle reading its 8
// corners here gets the bits it would get from the stencils at its corners,
// while the stencils run once per cell instead of 24 times per particle.
// Every thread computes its cells into registers; after a barrier (all
// potential reads done) it stores them over the potential window; a second
// barrier publishes them.  The launch has T threads.
template <int T = kWindowThreads>
__device__ __forceinline__ void build_force_window(float *lds) {
  constexpr int PER = (kWinCells + T - 1) / T;
  float f[PER][3];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e0 = int(threadIdx.x) + i * T;
    const int e = e0 < kWinCells ? e0 : kWinCells - 1;
    const int c = e % kWin, ab = e / kWin;
    const int b = ab % kWin, a = ab / kWin;
    stencil_lds(lds, a + kHalo, b + kHalo, c + kHalo, f[i]);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = int(threadIdx.x) + i * T;
    if (e < kWinCells) {
      lds[e] = f[i][0];
      lds[kWinCells + e] = f[i][1];
      lds[2 * kWinCells + e] = f[i][2];
    }
  }
  __syncthreads();
}

// The forces at a particle's 8 corners from the force window, corner o =
// bx | by << 1 | bz << 2 at window cell (a + bx, b + by, c + bz).
__device__ __forceinline__ void corners_fw(const float *fw, int a, int b, int c, float F[3][8]) {
  const float *p = fw + (a * kWin + b) * kWin + c;
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const int off = (o & 1) * kWin * kWin + ((o >> 1) & 1) * kWin + ((o >> 2) & 1);
    F[0][o] = p[off];
    F[1][o] = p[kWinCells + off];
    F[2][o] = p[2 * kWinCells + off];
  }
}

// -D phi at one window cell from the force window.
__device__ __forceinline__ void force_fw(const float *fw, int a, int b, int c, float F[3]) {
  const int e = (a * kWin + b) * kWin + c;
  F[0] = fw[e];
  F[1] = fw[kWinCells + e];
  F[2] = fw[2 * kWinCells + e];
}

// The 8 corner values of a mesh around an absolute position (cells floor(x)
// and floor(x) + 1 per axis, periodic), corner o = bx | by << 1 | bz << 2,
// read from global memory.  A floor cell already inside the mesh skips the
// periodic wrap's division.
__device__ __forceinline__ void corners_global_abs(const float *__restrict__ mesh, const TileGrid &t, float4 r,
                                                   float out[8]) {
  const float p[3] = {r.x, r.y, r.z};
  const int n[3] = {t.nx, t.ny, t.nz};
  uint32_t lo[3], hi[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int fi = cell_int(floorf(p[a]));
    const int c = uint32_t(fi) < uint32_t(n[a]) ? fi : wrapi(fi, n[a]);
    lo[a] = uint32_t(c);
    hi[a] = c + 1 == n[a] ? 0u : uint32_t(c + 1);
  }
  const uint32_t ny = uint32_t(t.ny), nz = uint32_t(t.nz);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const uint32_t x = (q & 1) ? hi[0] : lo[0], y = (q & 2) ? hi[1] : lo[1], z = (q & 4) ? hi[2] : lo[2];
    out[q] = mesh[(x * ny + y) * nz + z];
  }
}

// The particle's 8 corners form one 2^3 block inside the window (always for
// absolute positions; relative ones can wrap or round outside).
__device__ __forceinline__ bool corners_in_window(const CornerSet &cs) {
  bool ok = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) ok = ok && cs.l[d][0] >= 0 && cs.l[d][1] == cs.l[d][0] + 1;
  return ok;
}

// Same stencil read straight from the mesh (out-of-window fallback).
__device__ __forceinline__ void stencil_global(const float *__restrict__ phi, const TileGrid &t, int x, int y, int z,
                                               float F[3]) {
  auto at = [&](int i, int j, int k) {
    return phi[(uint32_t(wrapn(i, t.nx)) * uint32_t(t.ny) + uint32_t(wrapn(j, t.ny))) * uint32_t(t.nz) +
               uint32_t(wrapn(k, t.nz))];
  };
  F[0] = -(kC8 * (at(x + 1, y, z) - at(x - 1, y, z)) - kC1 * (at(x + 2, y, z) - at(x - 2, y, z)));
  F[1] = -(kC8 * (at(x, y + 1, z) - at(x, y - 1, z)) - kC1 * (at(x, y + 2, z) - at(x, y - 2, z)));
  F[2] = -(kC8 * (at(x, y, z + 1) - at(x, y, z - 1)) - kC1 * (at(x, y, z + 2) - at(x, y, z - 2)));
}

// The leapfrog drift that may follow a kick, fused into it: x = x + beta v
// with the new velocity (beta per simulation, beta_dev[sim], when given),
// written to xout and, when given, to xkeep.  next_count: NULL, or the tile
// counts (zeroed) that the drifted positions add to, for the next sort of
// them (absolute positions).
struct DriftArgs {
  float beta;
  const float *beta_dev;
  float *xout, *xkeep;
  uint32_t *next_count;
  float *vout;   // the kicked velocities (MODE 0 and 2); null: in place
};

// Tile of the floor cell of an absolute position, per axis, and its offset
// from tile (itx, ity, itz) folded periodically into {-1, 0, 1} when the two
// tiles are neighbours.  Returns whether all three offsets are.
__device__ __forceinline__ bool tile_offset(float x, float y, float z, const TileGrid &t, int itx, int ity, int itz,
                                            int &jx, int &jy, int &jz, int &d) {
  jx = int(uint32_t(floor_cell_abs(x, t.nx)) / uint32_t(kTile));
  jy = int(uint32_t(floor_cell_abs(y, t.ny)) / uint32_t(kTile));
  jz = int(uint32_t(floor_cell_abs(z, t.nz)) / uint32_t(kTile));
  const int j[3] = {jx, jy, jz}, i[3] = {itx, ity, itz}, T[3] = {t.tx, t.ty, t.tz};
  bool near = true;
  d = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int o = j[a] - i[a];
    if (o > 1) o -= T[a];
    if (o < -1) o += T[a];
    near = near && o >= -1 && o <= 1;
    d = d * 3 + (o + 1);
  }
  return near;
}

// MODE 0: vel[idx] += gamma * F;  MODE 1: out[idx] = gamma * F;  MODE 2:
// MODE 0 followed by the drift x[idx] = x + beta vel[idx], x being the
// sorted copy of the positions the forces were computed at.  gamma is per
// simulation (gamma_dev[sim]) when gamma_dev is given.
// Threads of a kick workgroup, two workgroups per CU (59 KiB of LDS each):
// 1024 (16 waves, at most 64 registers a thread) for absolute positions,
// 768 (12 waves, at most 80 registers) for relative ones, whose corner
// maths needs more registers.
template <bool REL>
constexpr int kick_threads() { return REL ? 768 : 1024; }

template <bool REL, int MODE, class G>
__global__ void __launch_bounds__(kick_threads<REL>(), REL ? 6 : 8)
bin_kick_kernel(BinsT<G> bins, const float *__restrict__ phi_all, float gamma_s, const float *__restrict__ gamma_dev,
                float *__restrict__ vel, uint32_t per, DriftArgs dr) {
  __shared__ float win[3 * kWinCells];   // the potential window, then the force window
  __shared__ uint32_t ncount[27];   // drifted particles per neighbour tile
  const uint32_t c = gather_item(blockIdx.x, per, bins.slabs);
  if (c >= *bins.total_chunks) return;
  const bool count_next = MODE == 2 && !REL && dr.next_count;
  if (count_next && threadIdx.x < 27) ncount[threadIdx.x] = 0u;
  const ChunkDesc d = bins.desc[c];
  const SimRef &sim = bins.g.tile_sim(d.tile);
  const TileGrid &t = sim.t;
  const uint32_t part0 = sim.part0;
  const float *phi = phi_all + sim.cell0;
  const float gamma = gamma_dev ? gamma_dev[sim.idx] : gamma_s;
  const float beta = MODE == 2 ? (dr.beta_dev ? dr.beta_dev[sim.idx] : dr.beta) : 0.0f;
  int itx, ity, itz;
  tile_origin(d.tile - sim.tile0, t, itx, ity, itz);
  const int ox = itx * kTile, oy = ity * kTile, oz = itz * kTile;
  // Software pipeline over this thread's particles q, q + B, q + 2B, ...:
  // an iteration computes particle q (record and velocity already loaded),
  // loads the velocity of q + B (whose record arrived an iteration ago) and
  // the record of q + 2B, so no load waits on another within an iteration.
  const uint32_t cnt = d.count, B = blockDim.x;
  const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float4 rec1 = threadIdx.x < cnt ? load_global(bins.spos + (d.start + threadIdx.x)) : zero4;
  float4 rec2 = threadIdx.x + B < cnt ? load_global(bins.spos + (d.start + threadIdx.x + B)) : zero4;
  float nv0 = 0.0f, nv1 = 0.0f, nv2 = 0.0f;
  if (MODE != 1 && threadIdx.x < cnt) {
    const float *v = vel + 3 * size_t(__float_as_uint(rec1.w));
    nv0 = v[0]; nv1 = v[1]; nv2 = v[2];
  }
  load_window_t<kPw, kHalo, kick_threads<REL>()>(phi, t, ox, oy, oz, win);
  __syncthreads();
  build_force_window<kick_threads<REL>()>(win);
  for (uint32_t q = threadIdx.x; q < cnt; q += B) {
    const float4 rec = rec1;
    const float pv0 = nv0, pv1 = nv1, pv2 = nv2;
    rec1 = rec2;
    if (MODE != 1 && q + B < cnt) {
      const float *v = vel + 3 * size_t(__float_as_uint(rec1.w));
      nv0 = v[0]; nv1 = v[1]; nv2 = v[2];
    }
    if (q + 2 * B < cnt) rec2 = load_global(bins.spos + (d.start + q + 2 * B));
    CornerSet cs;
    corner_set<REL>(rec, t, part0, ox, oy, oz, cs);
    float s[3] = {0.0f, 0.0f, 0.0f};
    if (!REL || corners_in_window(cs)) {
      float F8[3][8];
      corners_fw(win, cs.l[0][0], cs.l[1][0], cs.l[2][0], F8);
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        const int bx = o & 1, by = (o >> 1) & 1, bz = (o >> 2) & 1;
        const float wc = cs.w[0][bx] * cs.w[1][by] * cs.w[2][bz];
        s[0] += F8[0][o] * wc;
        s[1] += F8[1][o] * wc;
        s[2] += F8[2][o] * wc;
      }
    } else {
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        const int bx = o & 1, by = (o >> 1) & 1, bz = (o >> 2) & 1;
        const int la = cs.l[0][bx], lb = cs.l[1][by], lc = cs.l[2][bz];
        if (la == -2 || lb == -2 || lc == -2) continue;
        const float wc = cs.w[0][bx] * cs.w[1][by] * cs.w[2][bz];
        float F[3];
        if (la >= 0 && lb >= 0 && lc >= 0)
          force_fw(win, la, lb, lc, F);
        else
          stencil_global(phi, t, cs.g[0][bx], cs.g[1][by], cs.g[2][bz], F);
        s[0] += F[0] * wc;
        s[1] += F[1] * wc;
        s[2] += F[2] * wc;
      }
    }
    const uint32_t idx = __float_as_uint(rec.w);
    float *v = (dr.vout ? dr.vout : vel) + 3 * size_t(idx);
    if (MODE == 0) {
      v[0] = pv0 + gamma * s[0];
      v[1] = pv1 + gamma * s[1];
      v[2] = pv2 + gamma * s[2];
    } else if (MODE == 2) {
      const float v0 = pv0 + gamma * s[0], v1 = pv1 + gamma * s[1], v2 = pv2 + gamma * s[2];
      v[0] = v0;
      v[1] = v1;
      v[2] = v2;
      const float x0 = rec.x + beta * v0, x1 = rec.y + beta * v1, x2 = rec.z + beta * v2;
      float *xo = dr.xout + 3 * size_t(idx);
      xo[0] = x0;
      xo[1] = x1;
      xo[2] = x2;
      if (dr.xkeep) {
        float *xk = dr.xkeep + 3 * size_t(idx);
        xk[0] = x0;
        xk[1] = x1;
        xk[2] = x2;
      }
      if (count_next) {
        // most particles stay in their tile: those add up across the
        // wavefront and one lane adds them, the others add one each
        int jx, jy, jz, d;
        const bool nb = tile_offset(x0, x1, x2, t, itx, ity, itz, jx, jy, jz, d);
        const bool own = nb && d == 13;
        const unsigned long long act = __ballot(1), mown = __ballot(own);
        if (int(threadIdx.x & 63) == __ffsll((long long)act) - 1 && mown)
          atomicAdd(&ncount[13], uint32_t(__popcll(mown)));
        if (!own) {
          if (nb)
            atomicAdd(&ncount[d], 1u);
          else
            atomicAdd(&dr.next_count[sim.tile0 + (uint32_t(jx) * uint32_t(t.ty) + uint32_t(jy)) * uint32_t(t.tz) +
                                     uint32_t(jz)], 1u);
        }
      }
    } else {
      v[0] = gamma * s[0];
      v[1] = gamma * s[1];
      v[2] = gamma * s[2];
    }
  }
  if (count_next) {
    __syncthreads();
    if (threadIdx.x < 27 && ncount[threadIdx.x]) {
      const int o = int(threadIdx.x);
      int jx = itx + o / 9 - 1, jy = ity + (o / 3) % 3 - 1, jz = itz + o % 3 - 1;
      jx = jx < 0 ? jx + t.tx : (jx >= t.tx ? jx - t.tx : jx);
      jy = jy < 0 ? jy + t.ty : (jy >= t.ty ? jy - t.ty : jy);
      jz = jz < 0 ? jz + t.tz : (jz >= t.tz ? jz - t.tz : jz);
      atomicAdd(&dr.next_count[sim.tile0 + (uint32_t(jx) * uint32_t(t.ty) + uint32_t(jy)) * uint32_t(t.tz) +
                               uint32_t(jz)],
                ncount[threadIdx.x]);
    }
  }
}

// gpos[idx] += scale * sum_d cot_d dread(F_d)/dx + dread(rho_bar)/dx, scale
// per simulation (scale_dev[sim]) when scale_dev is given.
// RLDS: the density-adjoint window in LDS (true) or read from the mesh
// through the caches (false: 37 KiB of LDS per workgroup instead of 77 KiB).
// Both forms hold two workgroups per CU (the cache form by its registers);
// the LDS window, which moves a fifth fewer bytes and runs in 768-thread
// workgroups, is usually the faster.
// With cot_out (the same array as cot, so neither is restrict), the adjoint
// of the drift that preceded the kick follows: cot[idx] += beta gpos[idx]
// with the new gpos (beta per simulation, beta_dev[sim], when given).
// Threads of a force-VJP workgroup, two workgroups per CU: 768 (12 waves,
// at most 80 registers a thread) for absolute positions with the LDS density
// window, 512 for the other forms, whose registers would not fit.
template <bool REL, bool RLDS>
constexpr int grad_threads() { return !REL && RLDS ? 768 : 512; }

template <bool REL, class G, bool RLDS>
__global__ void __launch_bounds__((grad_threads<REL, RLDS>()), ((grad_threads<REL, RLDS>()) == 768 ? 6 : 1))
bin_grad_kernel(BinsT<G> bins, const float *__restrict__ phi_all, const float *__restrict__ rho_all,
                const float *cot, float scale_s, const float *__restrict__ scale_dev,
                float *__restrict__ gpos, uint32_t per, float beta_s, const float *__restrict__ beta_dev,
                float *cot_out, float *__restrict__ cot_max) {
  __shared__ float win[3 * kWinCells];   // the potential window, then the force window
  __shared__ float rwin[RLDS ? kWinCells : 1];
  constexpr int T = grad_threads<REL, RLDS>();
  __shared__ float wmax[T / 64];
  const uint32_t c = gather_item(blockIdx.x, per, bins.slabs);
  if (c >= *bins.total_chunks) return;
  const ChunkDesc d = bins.desc[c];
  const SimRef &sim = bins.g.tile_sim(d.tile);
  const TileGrid &t = sim.t;
  const uint32_t part0 = sim.part0;
  const float *phi = phi_all + sim.cell0;
  const float *rho_bar = rho_all + sim.cell0;
  const float scale = scale_dev ? scale_dev[sim.idx] : scale_s;
  const float beta = beta_dev ? beta_dev[sim.idx] : beta_s;
  int itx, ity, itz;
  tile_origin(d.tile - sim.tile0, t, itx, ity, itz);
  const int ox = itx * kTile, oy = ity * kTile, oz = itz * kTile;
  // Software pipeline over this thread's particles q, q + B, q + 2B, ...:
  // particle q computes while the cotangent and gpos of q + B (whose record
  // arrived an iteration ago) and the record of q + 2B load, so no load waits
  // on another within an iteration.  With absolute positions and the density
  // adjoint read through the caches (PF), the 8 density-adjoint corners of
  // q + B load in the same iteration, a particle ahead as well.
  constexpr bool PF = !REL && !RLDS;
  const uint32_t cnt = d.count, B = blockDim.x;
  const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float4 rec1 = threadIdx.x < cnt ? load_global(bins.spos + (d.start + threadIdx.x)) : zero4;
  float4 rec2 = threadIdx.x + B < cnt ? load_global(bins.spos + (d.start + threadIdx.x + B)) : zero4;
  float nu[3] = {0.0f, 0.0f, 0.0f}, ng[3] = {0.0f, 0.0f, 0.0f};
  float nr[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (threadIdx.x < cnt) {
    const size_t i3 = 3 * size_t(__float_as_uint(rec1.w));
    nu[0] = cot[i3]; nu[1] = cot[i3 + 1]; nu[2] = cot[i3 + 2];
    ng[0] = gpos[i3]; ng[1] = gpos[i3 + 1]; ng[2] = gpos[i3 + 2];
    if (PF) corners_global_abs(rho_bar, t, rec1, nr);
  }
  const bool near = near_wrap_ok(t);
  if (!RLDS)
    load_window_t<kPw, kHalo, T>(phi, t, ox, oy, oz, win);
  else if (near)
    load_window_pair<kPw, kHalo, kWin, 0, true, T>(phi, win, rho_bar, rwin, t, ox, oy, oz);
  else
    load_window_pair<kPw, kHalo, kWin, 0, false, T>(phi, win, rho_bar, rwin, t, ox, oy, oz);
  float cmax = 0.0f;
  __syncthreads();
  build_force_window<T>(win);
  for (uint32_t q = threadIdx.x; q < cnt; q += B) {
    const float4 rec = rec1;
    const uint32_t idx = __float_as_uint(rec.w);
    const float u0 = nu[0], u1 = nu[1], u2 = nu[2];
    const float g0 = ng[0], g1 = ng[1], g2 = ng[2];
    float r8[8];
#pragma unroll
    for (int o = 0; o < 8; ++o) r8[o] = nr[o];
    rec1 = rec2;
    if (q + B < cnt) {
      const size_t i3 = 3 * size_t(__float_as_uint(rec1.w));
      nu[0] = cot[i3]; nu[1] = cot[i3 + 1]; nu[2] = cot[i3 + 2];
      ng[0] = gpos[i3]; ng[1] = gpos[i3 + 1]; ng[2] = gpos[i3 + 2];
      if (PF) corners_global_abs(rho_bar, t, rec1, nr);
    }
    if (q + 2 * B < cnt) rec2 = load_global(bins.spos + (d.start + q + 2 * B));
    CornerSet cs;
    corner_set<REL>(rec, t, part0, ox, oy, oz, cs);
    float gd[3] = {0.0f, 0.0f, 0.0f}, gr[3] = {0.0f, 0.0f, 0.0f};
    if (!REL || corners_in_window(cs)) {
      float f8[8];
      const int la0 = cs.l[0][0], lb0 = cs.l[1][0], lc0 = cs.l[2][0];
      {
        // u . (-D phi) at the 8 corners, summed as (u0 Fx + u1 Fy) + u2 Fz
        float F8[3][8];
        corners_fw(win, la0, lb0, lc0, F8);
#pragma unroll
        for (int o = 0; o < 8; ++o) f8[o] = u0 * F8[0][o];
#pragma unroll
        for (int o = 0; o < 8; ++o) f8[o] = f8[o] + u1 * F8[1][o];
#pragma unroll
        for (int o = 0; o < 8; ++o) f8[o] = f8[o] + u2 * F8[2][o];
      }
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        const int bx = o & 1, by = (o >> 1) & 1, bz = (o >> 2) & 1;
        float rv;
        if (PF) {
          rv = r8[o];
        } else if (RLDS) {
          rv = rwin[((la0 + bx) * kWin + lb0 + by) * kWin + lc0 + bz];
        } else if (near) {
          const int x = wrap_near(ox + la0 + bx, t.nx), y = wrap_near(oy + lb0 + by, t.ny),
                    z = wrap_near(oz + lc0 + bz, t.nz);
          rv = rho_bar[(uint32_t(x) * uint32_t(t.ny) + uint32_t(y)) * uint32_t(t.nz) + uint32_t(z)];
        } else {
          const int x = wrapn(ox + la0 + bx, t.nx), y = wrapn(oy + lb0 + by, t.ny), z = wrapn(oz + lc0 + bz, t.nz);
          rv = rho_bar[(uint32_t(x) * uint32_t(t.ny) + uint32_t(y)) * uint32_t(t.nz) + uint32_t(z)];
        }
        const float f = f8[o];
        const float dwx = cs.d[0][bx] * cs.w[1][by] * cs.w[2][bz];
        const float dwy = cs.w[0][bx] * cs.d[1][by] * cs.w[2][bz];
        const float dwz = cs.w[0][bx] * cs.w[1][by] * cs.d[2][bz];
        gd[0] += f * dwx;
        gd[1] += f * dwy;
        gd[2] += f * dwz;
        gr[0] += rv * dwx;
        gr[1] += rv * dwy;
        gr[2] += rv * dwz;
      }
    } else {
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        const int bx = o & 1, by = (o >> 1) & 1, bz = (o >> 2) & 1;
        const int la = cs.l[0][bx], lb = cs.l[1][by], lc = cs.l[2][bz];
        if (la == -2 || lb == -2 || lc == -2) continue;
        float F[3], rv;
        if (la >= 0 && lb >= 0 && lc >= 0) {
          force_fw(win, la, lb, lc, F);
          if (RLDS) {
            rv = rwin[(la * kWin + lb) * kWin + lc];
          } else {
            const int x = wrapn(ox + la, t.nx), y = wrapn(oy + lb, t.ny), z = wrapn(oz + lc, t.nz);
            rv = rho_bar[(uint32_t(x) * uint32_t(t.ny) + uint32_t(y)) * uint32_t(t.nz) + uint32_t(z)];
          }
        } else {
          const int x = cs.g[0][bx], y = cs.g[1][by], z = cs.g[2][bz];
          stencil_global(phi, t, x, y, z, F);
          rv = rho_bar[(uint32_t(x) * uint32_t(t.ny) + uint32_t(y)) * uint32_t(t.nz) + uint32_t(z)];
        }
        const float f = u0 * F[0] + u1 * F[1] + u2 * F[2];
        const float dwx = cs.d[0][bx] * cs.w[1][by] * cs.w[2][bz];
        const float dwy = cs.w[0][bx] * cs.d[1][by] * cs.w[2][bz];
        const float dwz = cs.w[0][bx] * cs.w[1][by] * cs.d[2][bz];
        gd[0] += f * dwx;
        gd[1] += f * dwy;
        gd[2] += f * dwz;
        gr[0] += rv * dwx;
        gr[1] += rv * dwy;
        gr[2] += rv * dwz;
      }
    }
    float *g = gpos + 3 * size_t(idx);
    const float n0 = g0 + (scale * gd[0] + gr[0]), n1 = g1 + (scale * gd[1] + gr[1]),
                n2 = g2 + (scale * gd[2] + gr[2]);
    g[0] = n0;
    g[1] = n1;
    g[2] = n2;
    if (cot_out) {
      float *co = cot_out + 3 * size_t(idx);
      const float c0 = u0 + beta * n0, c1 = u1 + beta * n1, c2 = u2 + beta * n2;
      co[0] = c0;
      co[1] = c1;
      co[2] = c2;
      cmax = fmaxf(cmax, fmaxf(finite_abs(c0), fmaxf(finite_abs(c1), finite_abs(c2))));
    }
  }
  if (cot_max) {
    // max |updated cotangent| for the next force VJP's fixed-point bound: one
    // atomic per workgroup; values are >= 0, so the unsigned bit pattern
    // orders like the float
    for (int off = 32; off > 0; off >>= 1) cmax = fmaxf(cmax, __shfl_xor(cmax, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = cmax;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < int(blockDim.x >> 6); ++w) cmax = fmaxf(cmax, wmax[w]);
      atomicMax(reinterpret_cast<unsigned int *>(cot_max), __float_as_uint(cmax));
    }
  }
}

// cic_read (painting.py:78-106) of nc meshes over the tile-sorted particles:
// each workgroup loads the tile's 17^3 window of every mesh into LDS and
// sums the 8 corners of its particles in painting.py's corner order
// (000,100,010,001,110,101,011,111), the order and float operations of the
// direct read, so both give the same bits.  Absolute positions only.
template <int NC, class G>
__global__ void __launch_bounds__(kWindowThreads)
bin_read_kernel(BinsT<G> bins, const float *__restrict__ mesh_all, uint32_t cstride, float *__restrict__ out,
                uint32_t per) {
  __shared__ float win[NC * kWinCells];
  const uint32_t c = gather_item(blockIdx.x, per, bins.slabs);
  if (c >= *bins.total_chunks) return;
  const ChunkDesc d = bins.desc[c];
  const SimRef &sim = bins.g.tile_sim(d.tile);
  const TileGrid &t = sim.t;
  const float *mesh = mesh_all + sim.cell0;
  int itx, ity, itz;
  tile_origin(d.tile - sim.tile0, t, itx, ity, itz);
  const int ox = itx * kTile, oy = ity * kTile, oz = itz * kTile;
#pragma unroll
  for (int ch = 0; ch < NC; ++ch) load_window_t<kWin, 0>(mesh + size_t(ch) * cstride, t, ox, oy, oz, win + ch * kWinCells);
  __syncthreads();
  const int xs[8] = {0, 1, 0, 0, 1, 1, 0, 1};
  const int ys[8] = {0, 0, 1, 0, 1, 0, 1, 1};
  const int zs[8] = {0, 0, 0, 1, 0, 1, 1, 1};
  for (uint32_t q = threadIdx.x; q < d.count; q += blockDim.x) {
    const float4 rec = load_global(bins.spos + (d.start + q));
    const LAxis ax = laxis_abs(rec.x, ox, t.nx), ay = laxis_abs(rec.y, oy, t.ny), az = laxis_abs(rec.z, oz, t.nz);
    int li[8];
    float w[8];
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      li[o] = ((xs[o] ? ax.l1 : ax.l0) * kWin + (ys[o] ? ay.l1 : ay.l0)) * kWin + (zs[o] ? az.l1 : az.l0);
      w[o] = (xs[o] ? ax.w1 : ax.w0) * (ys[o] ? ay.w1 : ay.w0) * (zs[o] ? az.w1 : az.w0);
    }
    const uint32_t idx = __float_as_uint(rec.w);
#pragma unroll
    for (int ch = 0; ch < NC; ++ch) {
      float sum = 0.0f;
#pragma unroll
      for (int o = 0; o < 8; ++o) sum += win[ch * kWinCells + li[o]] * w[o];
      out[size_t(idx) * NC + ch] = sum;
    }
  }
}

template <class G>
static int kick_sorted(hipStream_t s, const BinsT<G> &b, const float *phi, float gamma, const float *gamma_dev,
                       float *vel, int mode, DriftArgs dr) {
  JPM_REQUIRE(phi && vel, "null pointer argument");
  JPM_REQUIRE(mode != 2 || dr.xout, "the fused drift needs its position output");
  JPM_REQUIRE(!dr.next_count || (mode == 2 && !b.rel), "tile counts of drifted absolute positions only");
  if (dr.next_count) {
    int rc = zero_u32(s, dr.next_count, b.K);
    if (rc) return rc;
  }
  uint32_t per;
  const dim3 grid(gather_grid(b.maxchunks, b.slabs, per)), block(b.rel ? kick_threads<true>() : kick_threads<false>());
  JPM_TIMED(s, mode == 2 ? "bin_kick_kernel[drift]" : mode == 1 ? "bin_kick_kernel[forces]" : "bin_kick_kernel");
  auto kern = b.rel ? (mode == 1 ? bin_kick_kernel<true, 1, G> : mode == 2 ? bin_kick_kernel<true, 2, G>
                                                                          : bin_kick_kernel<true, 0, G>)
                    : (mode == 1 ? bin_kick_kernel<false, 1, G> : mode == 2 ? bin_kick_kernel<false, 2, G>
                                                                           : bin_kick_kernel<false, 0, G>);
  hipLaunchKernelGGL(kern, grid, block, 0, s, b, phi, gamma, gamma_dev, vel, per, dr);
  JPM_LAUNCHED("bin_kick_kernel");
  return JPM_OK;
}

int zero_f32(hipStream_t s, float *p, int64_t m);

template <class G>
static int grad_sorted(hipStream_t s, const BinsT<G> &b, const float *phi, const float *rho_bar, const float *cot,
                       float scale, const float *scale_dev, float *gpos, bool rho_lds, float beta,
                       const float *beta_dev, float *cot_out, float *cot_max) {
  JPM_REQUIRE(phi && rho_bar && cot && gpos, "null pointer argument");
  JPM_REQUIRE(!cot_out || cot_out == cot, "the fused drift adjoint updates the cotangent in place");
  JPM_REQUIRE(!cot_max || cot_out, "the cotangent maximum is of the updated cotangent");
  if (cot_max) {
    int rc = zero_f32(s, cot_max, 1);
    if (rc) return rc;
  }
  uint32_t per;
  const int threads = b.rel ? (rho_lds ? grad_threads<true, true>() : grad_threads<true, false>())
                            : (rho_lds ? grad_threads<false, true>() : grad_threads<false, false>());
  const dim3 grid(gather_grid(b.maxchunks, b.slabs, per)), block(threads);
  JPM_TIMED(s, cot_out ? "bin_grad_kernel[drift]" : "bin_grad_kernel");
  auto kern = b.rel ? (rho_lds ? bin_grad_kernel<true, G, true> : bin_grad_kernel<true, G, false>)
                    : (rho_lds ? bin_grad_kernel<false, G, true> : bin_grad_kernel<false, G, false>);
  hipLaunchKernelGGL(kern, grid, block, 0, s, b, phi, rho_bar, cot, scale, scale_dev, gpos, per, beta, beta_dev,
                     cot_out, cot_max);
  JPM_LAUNCHED("bin_grad_kernel");
  return JPM_OK;
}

int batch_kick(hipStream_t s, const BatchBins &b, const float *phi, const float *gamma_dev, float *vel,
               int write_forces, const float *beta_dev, float *xout, float *xkeep, uint32_t *next_count, float *vout) {
  const int mode = write_forces ? 1 : (xout ? 2 : 0);
  return kick_sorted(s, b, phi, 0.0f, gamma_dev, vel, mode,
                     DriftArgs{0.0f, beta_dev, xout, xkeep, next_count, write_forces ? nullptr : vout});
}

int batch_grad(hipStream_t s, const BatchBins &b, const float *phi, const float *rho_bar, const float *cot,
               const float *scale_dev, float *gpos, bool rho_lds, const float *beta_dev, float *cot_out,
               float *cot_max) {
  return grad_sorted(s, b, phi, rho_bar, cot, 0.0f, scale_dev, gpos, rho_lds, 0.0f, beta_dev, cot_out, cot_max);
}

int bins_kick(hipStream_t s, const Bins &b, const float *phi, float gamma, float *vel, int write_forces, float beta,
              float *xout, float *xkeep, uint32_t *next_count, float *vout) {
  const int mode = write_forces ? 1 : (xout ? 2 : 0);
  return kick_sorted(s, b, phi, gamma, nullptr, vel, mode,
                     DriftArgs{beta, nullptr, xout, xkeep, next_count, write_forces ? nullptr : vout});
}

int bins_grad(hipStream_t s, const Bins &b, const float *phi, const float *rho_bar, const float *cot, float scale,
              float *gpos, bool rho_lds, float beta, float *cot_out, float *cot_max) {
  return grad_sorted(s, b, phi, rho_bar, cot, scale, nullptr, gpos, rho_lds, beta, nullptr, cot_out, cot_max);
}

int bins_read(hipStream_t s, const Bins &b, const float *mesh, int nc, uint32_t cstride, float *out) {
  JPM_REQUIRE(mesh && out, "null pointer argument");
  JPM_REQUIRE(!b.rel, "the tile-sorted read takes absolute positions");
  JPM_REQUIRE(nc >= 1 && nc <= 3, "the tile-sorted read takes 1 to 3 meshes (got %d)", nc);
  uint32_t per;
  const dim3 grid(gather_grid(b.maxchunks, b.slabs, per)), block(kWindowThreads);
  JPM_TIMED(s, "bin_read_kernel");
  auto kern = nc == 1 ? bin_read_kernel<1, OneSim> : nc == 2 ? bin_read_kernel<2, OneSim> : bin_read_kernel<3, OneSim>;
  hipLaunchKernelGGL(kern, grid, block, 0, s, b, mesh, cstride, out, per);
  JPM_LAUNCHED("bin_read_kernel");
  return JPM_OK;
}

}  // namespace jpm
