// Batches of simulations of different mesh sizes and cosmologies, stepped
// together: every stage of a force evaluation (tile sort, window paint,
// potential, gather) is one launch over all simulations of the batch --
// kernels look their simulation up in a device table by tile or workgroup
// index -- except the FFTs and the per-shape k-space passes, which run once
// per mesh shape as strided batches.  Per-simulation leapfrog coefficients
// are device arrays, so a captured time loop replays without host work.
//
// Layout: the batch concatenates its simulations in order.  Particles
// [N][3], meshes [M], half spectra [H]; simulations of one shape are
// contiguous.
#include <algorithm>
#include <new>

#include "tiles.hpp"

namespace jpm {

int fft_r2c(hipStream_t, int, int, int, int, const float *, void *);
int fft_c2r(hipStream_t, int, int, int, int, void *, float *);
int kspace_potential_batch(hipStream_t, int, int, int, int, const void *, void *, float, float, const float *,
                           uint32_t);
int fd4_divergence_batch(hipStream_t, int, int, int, int, const float *, size_t, float, float *);
int batch_kick(hipStream_t, const BatchBins &, const float *, const float *, float *, int, const float *, float *,
               float *, uint32_t *, float *);
int kspace_force_batch(hipStream_t, int, int, int, int, const void *, void *, float, uint32_t);
int kspace_force_adjoint_batch(hipStream_t, int, int, int, int, const void *, void *, float, uint32_t);
int kspace_linear_field_batch(hipStream_t, int, int, int, int, void *, const float *, const float *, const float *, int,
                              float);
int grad_pos_batch(hipStream_t, const SimRef *, int, bool, const float *, const float *, float *);
int batch_grad(hipStream_t, const BatchBins &, const float *, const float *, const float *, const float *, float *,
               bool, const float *, float *, float *);

// ------------------------------------------------------------ description

int batch_create(int B, const int *dims, const int64_t *counts, uint32_t flags, SimBatch **out) {
  JPM_REQUIRE(out, "null output handle");
  *out = nullptr;
  JPM_REQUIRE(B >= 1 && B <= 65535, "batch of %d simulations", B);
  JPM_REQUIRE(dims && counts, "null dims or counts");
  const bool rel = flags & JPM_RELATIVE;
  SimBatch *bt = new (std::nothrow) SimBatch;
  JPM_REQUIRE(bt, "out of host memory");
  bt->B = B;
  bt->rel = rel;
  bt->chunk = tuned_chunk(flags);
  bt->slice = tuned_slice(flags);
  bt->rho_lds = tuned_rho_lds(flags);
  bt->slabs = tuned_slabs(flags);
  uint64_t K = 0, N = 0, M = 0, H = 0, G = 0;
  uint32_t maxK = 0;
  for (int b = 0; b < B; ++b) {
    const int nx = dims[3 * b], ny = dims[3 * b + 1], nz = dims[3 * b + 2];
    const int64_t n = counts[b];
    if (int rc = check_dims(nx, ny, nz)) {
      delete bt;
      return rc;
    }
    const TileGrid t = make_tiles(nx, ny, nz);
    const int64_t m = cells(nx, ny, nz);
    if (n < 1 || (rel && n != m)) {
      delete bt;
      return fail(JPM_E_INVALID, "simulation %d: mesh (%d,%d,%d) with %lld particles (%s)", b, nx, ny, nz,
                  (long long)n, rel ? "relative mode needs one particle per cell" : "no particles");
    }
    SimRef r;
    r.t = t;
    r.tile0 = uint32_t(K);
    r.part0 = uint32_t(N);
    r.n = uint32_t(n);
    r.cell0 = uint32_t(M);
    r.blk0 = uint32_t(G);
    r.idx = uint32_t(b);
    bt->host.push_back(r);
    SimClass *last = bt->classes.empty() ? nullptr : &bt->classes.back();
    if (last && last->nx == nx && last->ny == ny && last->nz == nz) {
      last->count++;
    } else {
      for (const SimClass &c : bt->classes)
        if (c.nx == nx && c.ny == ny && c.nz == nz) {
          delete bt;
          return fail(JPM_E_INVALID, "simulations of mesh (%d,%d,%d) must be contiguous in the batch", nx, ny, nz);
        }
      SimClass c;
      c.nx = nx;
      c.ny = ny;
      c.nz = nz;
      c.first = b;
      c.count = 1;
      c.cell0 = uint32_t(M);
      c.half0 = uint32_t(H);
      bt->classes.push_back(c);
    }
    K += t.K;
    N += uint64_t(n);
    M += uint64_t(m);
    H += uint64_t(half_cells(nx, ny, nz));
    G += (uint64_t(n) + bt->slice - 1) / bt->slice;
    maxK = std::max(maxK, t.K);
  }
  if (N > uint64_t(kMaxCells) || 3 * M > (uint64_t(1) << 32) || 3 * H > (uint64_t(1) << 32)) {
    delete bt;
    return fail(JPM_E_INVALID, "batch too large: %llu particles, %llu cells", (unsigned long long)N,
                (unsigned long long)M);
  }
  bt->K = uint32_t(K);
  bt->N = uint32_t(N);
  bt->M = uint32_t(M);
  bt->H = uint32_t(H);
  bt->G = uint32_t(G);
  bt->maxK = maxK;
  hipError_t e = hipMalloc(&bt->dev, sizeof(SimRef) * size_t(B));
  if (e == hipSuccess) e = hipMemcpy(bt->dev, bt->host.data(), sizeof(SimRef) * size_t(B), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (bt->dev) (void)hipFree(bt->dev);
    delete bt;
    return fail(JPM_E_HIP, "uploading the simulation table failed: %s", hipGetErrorString(e));
  }
  *out = bt;
  return JPM_OK;
}

int batch_destroy(SimBatch *bt) {
  if (!bt) return JPM_OK;
  hipError_t e = bt->dev ? hipFree(bt->dev) : hipSuccess;
  delete bt;
  if (e != hipSuccess) return fail(JPM_E_HIP, "freeing the simulation table failed: %s", hipGetErrorString(e));
  return JPM_OK;
}

// --------------------------------------------------------------- workspace

struct BatchWs {
  size_t sort, paint, mesh, g3, spec, spec2, total;
};

static BatchWs batch_ws(const SimBatch &bt, int nc) {
  BatchWs w;
  size_t o = 0;
  w.sort = o;  o += align_up(batch_sort_workspace_size(bt));
  w.paint = o; o += align_up(batch_paint_workspace_size(bt, nc));
  w.mesh = o;  o += align_up(size_t(bt.M) * 4);
  w.g3 = o;    o += nc == 3 ? align_up(3 * size_t(bt.M) * 4) : 0;
  w.spec = o;  o += align_up(size_t(bt.H) * 8);
  w.spec2 = o; o += align_up(size_t(bt.H) * 8);
  w.total = o;
  return w;
}

size_t batch_kick_ws(const SimBatch &bt) { return batch_ws(bt, 1).total; }
size_t batch_vjp_ws(const SimBatch &bt) { return batch_ws(bt, 3).total; }

// ---------------------------------------------------------------- pipelines

static int check_kflags(uint32_t kflags) {
  JPM_REQUIRE(!(kflags & JPM_K_GRAD_EXACT), "batched pipelines use the 4-point gradient (no JPM_K_GRAD_EXACT)");
  return JPM_OK;
}

// For every simulation b: phi_b = potential of the CIC density of its
// particles; vel_b += gamma[b] F_b(pos_b) (or = F_b with write_forces).  With
// xout, the drift follows: xout_b = pos_b + beta[b] vel_b (also into xkeep
// when given; xout may be pos).
int batch_force_kick(hipStream_t s, const SimBatch &bt, const float *pos, float *vel, const float *gamma, float *phi,
                     float r_split, uint32_t kflags, int write_forces, void *ws, size_t ws_bytes, void *keep,
                     size_t keep_bytes, const float *beta, float *xout, float *xkeep, const uint32_t *counts_in,
                     uint32_t *counts_out, float *vout) {
  JPM_REQUIRE(!xout || (beta && !write_forces), "the fused drift needs beta and a kick");
  JPM_REQUIRE(!counts_out || (xout && !bt.rel), "tile counts of drifted absolute positions only");
  if (int rc = check_kflags(kflags)) return rc;
  JPM_REQUIRE(pos && phi && (gamma || !vel), "null pointer argument");
  JPM_REQUIRE(vel || !(write_forces || xout), "the force write and the fused drift need vel");
  const BatchWs w = batch_ws(bt, 1);
  JPM_REQUIRE(ws && ws_bytes >= w.total, "batched force_kick needs %zu workspace bytes, got %zu", w.total, ws_bytes);
  char *base = static_cast<char *>(ws);
  BatchBins b;
  int rc;
  if (keep) {
    JPM_REQUIRE(keep_bytes >= batch_sort_workspace_size(bt), "keep_sort buffer too small (%zu bytes)", keep_bytes);
    rc = batch_sort(s, bt, pos, keep, keep_bytes, b, counts_in);
  } else {
    rc = batch_sort(s, bt, pos, base + w.sort, w.paint - w.sort, b, counts_in);
  }
  if (rc) return rc;
  float *mesh = reinterpret_cast<float *>(base + w.mesh);
  float2 *spec = reinterpret_cast<float2 *>(base + w.spec), *spec2 = reinterpret_cast<float2 *>(base + w.spec2);
  rc = batch_paint(s, bt, b, nullptr, 1, mesh, base + w.paint, w.mesh - w.paint);
  if (rc) return rc;
  for (const SimClass &c : bt.classes) {
    const float inv_m = float(1.0 / double(cells(c.nx, c.ny, c.nz)));
    if ((rc = potential_solve(s, c.nx, c.ny, c.nz, c.count, mesh + c.cell0, spec + c.half0, spec2 + c.half0,
                              phi + c.cell0, r_split, inv_m, nullptr, kflags)))
      return rc;
  }
  // without vel the call is the force evaluation alone: potentials and sort
  if (!vel) return JPM_OK;
  return batch_kick(s, b, phi, gamma, vel, write_forces, beta, xout, xkeep, counts_out, vout);
}

// For every simulation b: gpos_b += scale[b] J_b^T cot_b, with pot_scale[b] =
// scale[b] / M_b applied inside the potential.  With beta, the adjoint of
// the drift that preceded the kick follows: cot_b += beta[b] gpos_b.
int batch_force_vjp(hipStream_t s, const SimBatch &bt, const float *phi, const float *pos, const float *cot,
                    const float *scale, const float *pot_scale, float *gpos, float r_split, uint32_t kflags, void *ws,
                    size_t ws_bytes, const void *sorted, size_t sorted_bytes, const float *beta, float *cot_out,
                    const float *cot_max_in, float *cot_max_out) {
  if (int rc = check_kflags(kflags)) return rc;
  JPM_REQUIRE(phi && pos && cot && scale && pot_scale && gpos, "null pointer argument");
  const BatchWs w = batch_ws(bt, 3);
  JPM_REQUIRE(ws && ws_bytes >= w.total, "batched force_vjp needs %zu workspace bytes, got %zu", w.total, ws_bytes);
  char *base = static_cast<char *>(ws);
  BatchBins b;
  int rc;
  if (sorted) {
    JPM_REQUIRE(sorted_bytes >= batch_sort_workspace_size(bt), "sorted buffer too small (%zu bytes)", sorted_bytes);
    b = batch_view(bt, sorted);
  } else {
    rc = batch_sort(s, bt, pos, base + w.sort, w.paint - w.sort, b);
    if (rc) return rc;
  }
  float *g3 = reinterpret_cast<float *>(base + w.g3);
  float *rho = reinterpret_cast<float *>(base + w.mesh);
  float2 *spec = reinterpret_cast<float2 *>(base + w.spec), *spec2 = reinterpret_cast<float2 *>(base + w.spec2);
  rc = batch_paint(s, bt, b, cot, 3, g3, base + w.paint, w.mesh - w.paint, cot_max_in);
  if (rc) return rc;
  for (const SimClass &c : bt.classes) {
    if ((rc = potential_solve_div(s, c.nx, c.ny, c.nz, c.count, g3 + c.cell0, size_t(bt.M), rho + c.cell0,
                                  spec + c.half0, spec2 + c.half0, rho + c.cell0, r_split, 0.0f, pot_scale + c.first,
                                  kflags)))
      return rc;
  }
  return batch_grad(s, b, phi, rho, cot, scale, gpos, bt.rho_lds, beta, cot_out, cot_max_out);
}

// ---------------------------------------- initial conditions and final paint
//
// The stages of the forward model around the leapfrog, batched the same way:
// the Gaussian field of each simulation from its white noise and P(k) table
// (pm.py:129-144), the first-order LPT forces read at the particle lattice
// (pm.py:79-82; a CIC read at lattice nodes returns the node values, so the
// read is a transpose of the three force meshes) with their adjoints, and
// the CIC paint of the final positions with its position cotangent.

struct LptWs {
  size_t spec, spec3, f3, total;
};

static LptWs lpt_ws(const SimBatch &bt) {
  size_t maxH = 0, maxM = 0;
  for (const SimClass &c : bt.classes) {
    maxH = std::max(maxH, size_t(c.count) * size_t(half_cells(c.nx, c.ny, c.nz)));
    maxM = std::max(maxM, size_t(c.count) * size_t(cells(c.nx, c.ny, c.nz)));
  }
  LptWs w;
  size_t o = 0;
  w.spec = o;  o += align_up(size_t(bt.H) * 8);
  w.spec3 = o; o += align_up(3 * maxH * 8);
  w.f3 = o;    o += align_up(3 * maxM * 4);
  w.total = o;
  return w;
}

size_t batch_lpt_ws(const SimBatch &bt) { return lpt_ws(bt).total; }

int batch_linear_field(hipStream_t s, const SimBatch &bt, const float *noise, float *out, const float *boxes,
                       const float *ktab, const float *pktabs, int ntab, void *ws, size_t ws_bytes) {
  JPM_REQUIRE(noise && out && boxes && ktab && pktabs, "null pointer argument");
  const LptWs w = lpt_ws(bt);
  JPM_REQUIRE(ws && ws_bytes >= w.total, "batched linear field needs %zu workspace bytes", w.total);
  float2 *spec = reinterpret_cast<float2 *>(static_cast<char *>(ws) + w.spec);
  int rc;
  for (const SimClass &c : bt.classes) {
    const float inv_m = float(1.0 / double(cells(c.nx, c.ny, c.nz)));
    if ((rc = fft_r2c(s, c.nx, c.ny, c.nz, c.count, noise + c.cell0, spec + c.half0))) return rc;
    if ((rc = kspace_linear_field_batch(s, c.nx, c.ny, c.nz, c.count, spec + c.half0, boxes + 3 * size_t(c.first), ktab,
                                        pktabs + size_t(c.first) * ntab, ntab, inv_m)))
      return rc;
    if ((rc = fft_c2r(s, c.nx, c.ny, c.nz, c.count, spec + c.half0, out + c.cell0))) return rc;
  }
  return JPM_OK;
}

// psi[(cell0_b + c) * 3 + d] <-> f3[b][d][c] for the simulations of one class.
__global__ void __launch_bounds__(256)
interleave3_kernel(float *__restrict__ psi, float *__restrict__ f3, uint32_t M, uint32_t count, int to_psi) {
  const size_t total = size_t(count) * M;
  for (size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += size_t(gridDim.x) * blockDim.x) {
    const size_t b = e / M, c = e - b * M;
    float *f = f3 + 3 * b * M + c;
    float *p = psi + 3 * e;
    if (to_psi) {
      p[0] = f[0];
      p[1] = f[M];
      p[2] = f[2 * size_t(M)];
    } else {
      f[0] = p[0];
      f[M] = p[1];
      f[2 * size_t(M)] = p[2];
    }
  }
}

static int check_lattice(const SimBatch &bt) {
  for (const SimRef &r : bt.host)
    JPM_REQUIRE(int64_t(r.n) == int64_t(r.t.nx) * r.t.ny * r.t.nz,
                "lattice forces need one particle per cell in lattice order (simulation %u)", r.idx);
  return JPM_OK;
}

int batch_lattice_forces(hipStream_t s, const SimBatch &bt, const float *field, float *psi, void *ws,
                         size_t ws_bytes) {
  JPM_REQUIRE(field && psi, "null pointer argument");
  if (int rc = check_lattice(bt)) return rc;
  const LptWs w = lpt_ws(bt);
  JPM_REQUIRE(ws && ws_bytes >= w.total, "batched lattice forces need %zu workspace bytes", w.total);
  char *base = static_cast<char *>(ws);
  float2 *spec = reinterpret_cast<float2 *>(base + w.spec), *spec3 = reinterpret_cast<float2 *>(base + w.spec3);
  float *f3 = reinterpret_cast<float *>(base + w.f3);
  int rc;
  for (const SimClass &c : bt.classes) {
    const int64_t m = cells(c.nx, c.ny, c.nz);
    if ((rc = fft_r2c(s, c.nx, c.ny, c.nz, c.count, field + c.cell0, spec + c.half0))) return rc;
    if ((rc = kspace_force_batch(s, c.nx, c.ny, c.nz, c.count, spec + c.half0, spec3, float(1.0 / double(m)), 0)))
      return rc;
    if ((rc = fft_c2r(s, c.nx, c.ny, c.nz, 3 * c.count, spec3, f3))) return rc;
    hipLaunchKernelGGL(interleave3_kernel, dim3(grid_for(int64_t(c.count) * m, 256)), dim3(256), 0, s,
                       psi + 3 * size_t(c.cell0), f3, uint32_t(m), uint32_t(c.count), 1);
    JPM_LAUNCHED("interleave3_kernel");
  }
  return JPM_OK;
}

int batch_lattice_forces_vjp(hipStream_t s, const SimBatch &bt, const float *gpsi, float *gfield, void *ws,
                             size_t ws_bytes) {
  JPM_REQUIRE(gpsi && gfield, "null pointer argument");
  if (int rc = check_lattice(bt)) return rc;
  const LptWs w = lpt_ws(bt);
  JPM_REQUIRE(ws && ws_bytes >= w.total, "batched lattice forces need %zu workspace bytes", w.total);
  char *base = static_cast<char *>(ws);
  float2 *spec = reinterpret_cast<float2 *>(base + w.spec), *spec3 = reinterpret_cast<float2 *>(base + w.spec3);
  float *f3 = reinterpret_cast<float *>(base + w.f3);
  int rc;
  for (const SimClass &c : bt.classes) {
    const int64_t m = cells(c.nx, c.ny, c.nz);
    hipLaunchKernelGGL(interleave3_kernel, dim3(grid_for(int64_t(c.count) * m, 256)), dim3(256), 0, s,
                       const_cast<float *>(gpsi) + 3 * size_t(c.cell0), f3, uint32_t(m), uint32_t(c.count), 0);
    JPM_LAUNCHED("interleave3_kernel");
    if ((rc = fft_r2c(s, c.nx, c.ny, c.nz, 3 * c.count, f3, spec3))) return rc;
    if ((rc = kspace_force_adjoint_batch(s, c.nx, c.ny, c.nz, c.count, spec3, spec + c.half0,
                                         float(1.0 / double(m)), 0)))
      return rc;
    if ((rc = fft_c2r(s, c.nx, c.ny, c.nz, c.count, spec + c.half0, gfield + c.cell0))) return rc;
  }
  return JPM_OK;
}

// The tile order of an absolute-position batch: each simulation's slots stay
// within its own particle range, since its tiles follow the tiles of the
// simulations before it.
int batch_tile_order(hipStream_t s, const SimBatch &bt, const float *pos, uint32_t *order, void *ws,
                     size_t ws_bytes) {
  JPM_REQUIRE(!bt.rel, "the tile order takes absolute positions (relative ones are tied to their lattice index)");
  JPM_REQUIRE(pos && order, "null pointer argument");
  const size_t need = batch_sort_workspace_size(bt);
  JPM_REQUIRE(ws && ws_bytes >= need, "tile order needs %zu workspace bytes, got %zu", need, ws_bytes);
  BatchBins b;
  int rc = batch_sort(s, bt, pos, ws, ws_bytes, b);
  if (rc) return rc;
  return bins_order(s, b, order);
}

int batch_paint_mesh(hipStream_t s, const SimBatch &bt, const float *pos, float *mesh, void *ws, size_t ws_bytes) {
  JPM_REQUIRE(pos && mesh, "null pointer argument");
  const BatchWs w = batch_ws(bt, 1);
  JPM_REQUIRE(ws && ws_bytes >= w.total, "batched paint needs %zu workspace bytes, got %zu", w.total, ws_bytes);
  char *base = static_cast<char *>(ws);
  BatchBins b;
  int rc = batch_sort(s, bt, pos, base + w.sort, w.paint - w.sort, b);
  if (rc) return rc;
  return batch_paint(s, bt, b, nullptr, 1, mesh, base + w.paint, w.mesh - w.paint);
}

int batch_paint_vjp(hipStream_t s, const SimBatch &bt, const float *pos, const float *gmesh, float *gpos) {
  return grad_pos_batch(s, bt.dev, bt.B, bt.rel, gmesh, pos, gpos);
}

// ------------------------------------------------- per-simulation vectors

constexpr int kVecBlocks = 64;   // workgroups per simulation

// y_b += a[b] x_b over each simulation's 3 n_b floats.
__global__ void __launch_bounds__(256)
batch_axpy_kernel(const SimRef *__restrict__ sims, const float *__restrict__ a, const float *__restrict__ x,
                  float *__restrict__ y) {
  const SimRef &sim = sims[blockIdx.y];
  const float c = a[blockIdx.y];
  const size_t lo = 3 * size_t(sim.part0), hi = lo + 3 * size_t(sim.n);
  for (size_t i = lo + size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < hi; i += size_t(gridDim.x) * blockDim.x)
    y[i] += c * x[i];
}

__device__ __forceinline__ double bsum_d(double v) {
  __shared__ double part[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < int(blockDim.x >> 6); ++w) s += part[w];
  return s;
}

__global__ void __launch_bounds__(256)
batch_dot_partial_kernel(const SimRef *__restrict__ sims, const float *__restrict__ x, const float *__restrict__ y,
                         double *__restrict__ partial) {
  const SimRef &sim = sims[blockIdx.y];
  const size_t lo = 3 * size_t(sim.part0), hi = lo + 3 * size_t(sim.n);
  double s = 0.0;
  for (size_t i = lo + size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < hi; i += size_t(gridDim.x) * blockDim.x)
    s += double(x[i]) * double(y[i]);
  s = bsum_d(s);
  if (threadIdx.x == 0) partial[size_t(blockIdx.y) * gridDim.x + blockIdx.x] = s;
}

__global__ void __launch_bounds__(256)
batch_dot_final_kernel(const double *__restrict__ partial, int blocks, double *__restrict__ out) {
  double s = 0.0;
  for (int b = threadIdx.x; b < blocks; b += blockDim.x) s += partial[size_t(blockIdx.x) * blocks + b];
  s = bsum_d(s);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

int batch_axpy(hipStream_t s, const SimBatch &bt, const float *a, const float *x, float *y) {
  JPM_REQUIRE(a && x && y, "null pointer argument");
  JPM_TIMED(s, "batch_axpy_kernel");
  hipLaunchKernelGGL(batch_axpy_kernel, dim3(kVecBlocks * 4, bt.B), dim3(256), 0, s, bt.dev, a, x, y);
  JPM_LAUNCHED("batch_axpy_kernel");
  return JPM_OK;
}

size_t batch_dot_ws(const SimBatch &bt) { return size_t(bt.B) * kVecBlocks * sizeof(double); }

int batch_dot(hipStream_t s, const SimBatch &bt, const float *x, const float *y, double *out, void *ws,
              size_t ws_bytes) {
  JPM_REQUIRE(x && y && out, "null pointer argument");
  JPM_REQUIRE(ws && ws_bytes >= batch_dot_ws(bt), "batched dot needs %zu workspace bytes", batch_dot_ws(bt));
  double *partial = static_cast<double *>(ws);
  hipLaunchKernelGGL(batch_dot_partial_kernel, dim3(kVecBlocks, bt.B), dim3(256), 0, s, bt.dev, x, y, partial);
  JPM_LAUNCHED("batch_dot_partial_kernel");
  hipLaunchKernelGGL(batch_dot_final_kernel, dim3(bt.B), dim3(256), 0, s, partial, kVecBlocks, out);
  JPM_LAUNCHED("batch_dot_final_kernel");
  return JPM_OK;
}

}  // namespace jpm
