/*
 * jaxpm_hip.h -- C ABI of libjaxpm_hip.so, the MI355X (gfx950) particle-mesh
 * step behind the jaxpm_amd Python package.
 *
 * The reference (JaxPM) is pure Python over JAX/XLA; it has no FFI of its own.
 * Each entry point below replaces the XLA computation that a reference
 * function lowers to on one device, cited as reference file:line.  A JAX
 * binding would wrap these as XLA-FFI custom calls (one handler per entry
 * point); the shipped binding is ctypes from jaxpm_amd/_lib.py.
 *
 * Conventions
 *   - Every pointer is device memory owned by the caller, except where noted.
 *   - `stream` is a hipStream_t passed as void* (0 = default stream).  All
 *     work is stream-ordered; no entry point synchronises the host unless
 *     its comment says so.
 *   - Mesh layout: C order [nx][ny][nz] float32 (z fastest), as the
 *     reference's jnp arrays.  Particle arrays: [n][3] float32 in grid units
 *     (cell = 1).  Half spectra: C order [nx][ny][nz/2+1] complex64
 *     (interleaved re, im), the hipFFT R2C layout.
 *   - Return value: 0 on success, a negative JPM_E* code on failure; the
 *     message of the last failure on the calling thread is jpm_last_error().
 *     No C++ exception crosses the ABI.
 *   - Re-entrant; the FFT plan and wave-number table caches are mutex-guarded.
 */
#ifndef JAXPM_HIP_H
#define JAXPM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JPM_ABI_VERSION 1

#define JPM_OK 0
#define JPM_E_INVALID -1   /* bad shape, size, flag or null pointer */
#define JPM_E_HIP -2       /* HIP runtime error */
#define JPM_E_FFT -3       /* hipFFT error */
#define JPM_E_NOMEM -4     /* workspace too small or allocation failed */

/* paint / read flags */
#define JPM_RELATIVE 1u      /* positions are displacements from the particle
                                lattice (i,j,k) of shape (nx,ny,nz), n == nx*ny*nz
                                (painting_utils.enmesh semantics) */
#define JPM_DETERMINISTIC 2u /* bitwise-reproducible paint (fixed-point sums) */
#define JPM_ZERO_MESH 4u     /* clear the output mesh before painting */
#define JPM_ATOMIC 8u        /* pipelines: paint with global atomics instead
                                of the binned LDS paint */

/* Pipeline flags (jpm_pm_forces_f32, jpm_force_*, jpm_batch_create): paint
 * flags in bits 0..7, k-space flags shifted left by 8, and the tile sort's
 * tuning in bits 16..26 -- JPM_TUNE_CHUNK(c): at most 2^(9+c) particles per
 * accumulation/gather chunk (c = 1..6, default 8192), JPM_TUNE_SLICE(s):
 * 2^(11+s) particles per counting/scatter workgroup (s = 1..7, default
 * 8192).  The choice changes speed, never results beyond float rounding;
 * a kept sort must be read with the flags it was built with. */
#define JPM_TUNE_CHUNK(c) ((uint32_t)(c) << 16)
#define JPM_TUNE_SLICE(s) ((uint32_t)(s) << 20)
#define JPM_TUNE_RHO_GLOBAL (1u << 24)  /* force-VJP gather: density adjoint
                                           read through the caches, not LDS */
#define JPM_TUNE_XCD_SLABS (1u << 25)   /* tile workgroups placed so each XCD
                                           walks one contiguous run of tiles */
#define JPM_TUNE_FACES_IN_MESH (1u << 26) /* jpm_force_kick*: the paint writes
                                           its tiles' low-face cells into the
                                           mesh instead of a buffer of their
                                           own read by the potential solve
                                           (absolute positions on meshes of
                                           full 16^3 tiles); same bits */

/* k-space flags */
#define JPM_K_GRAD_EXACT 1u  /* gradient_kernel order=0 (i k, len//2 zeroed)
                                instead of the default 4-point order=1 */
#define JPM_K_LAPL_FD 2u     /* invlaplace_kernel(fd=True) */
#define JPM_K_DECONV 4u      /* multiply by cic_compensation (opt-in) */

/* Library lifetime and diagnostics. */
int jpm_init(void);
int jpm_shutdown(void);                 /* frees cached FFT plans and tables */
int jpm_abi_version(void);
const char *jpm_last_error(void);

/* Per-kernel timing for benchmarks: while enabled, the library records a
 * pair of HIP events on the launch stream around each of its kernels;
 * jpm_timing_report waits for them and writes "name launches total_ms" lines
 * (then clears).  No counterpart in the reference. */
int jpm_timing_enable(int on);
int jpm_timing_report(char *buf, size_t len);

/* ---------------------------------------------------------------- painting */

/* Non-finite input.  Every paint gives a deposit that is NaN or +-Inf (a
 * non-finite weight, or a NaN / +-Inf position, whose corner weights are
 * NaN) to the cells it touches as a float, as the reference's float
 * scatter_add does: a touched cell becomes NaN, +Inf or -Inf (NaN where
 * both infinities land).  The fixed-point sums of the deterministic paints
 * cover the finite deposits only, scaled by the largest finite weight.  A
 * non-finite position indexes the cells its floor would if it were 0
 * (absolute mode; the reference's int32 cast of it is implementation-defined)
 * and relative mode drops it, as enmesh's mode='drop' scatter does. */

/* Bytes of workspace jpm_cic_paint_f32 needs for this call (0 unless
 * JPM_DETERMINISTIC: 8 bytes per mesh cell). */
size_t jpm_cic_paint_workspace_size(int nx, int ny, int nz, int64_t n,
                                    uint32_t flags);

/* CIC deposit of n particles into mesh[nx][ny][nz] (accumulates unless
 * JPM_ZERO_MESH).  weight: per-particle weights or NULL for the scalar
 * `wscalar`.
 * Replaces painting.py:15-45 (_cic_paint_impl: 8 corners, weight
 * prod(1-|x-c|), int32 cast then periodic mod, scatter_add) and, with
 * JPM_RELATIVE, painting.py:161-189 + painting_utils.py:28-131 (enmesh +
 * chunked scatter; corners that round onto index L are dropped). */
int jpm_cic_paint_f32(void *stream, int nx, int ny, int nz, const float *pos,
                      int64_t n, const float *weight, float wscalar,
                      float *mesh, uint32_t flags, void *ws, size_t ws_bytes);

/* CIC gather of `ncomp` (1..4) meshes stored `comp_stride` floats apart:
 * out[p*ncomp + c] = sum_o mesh_c[idx_o(p)] W_o(p).
 * Replaces painting.py:78-106 (_cic_read_impl) and, with JPM_RELATIVE,
 * painting.py:218-236 + painting_utils.py:144-187 (mode='drop' gather);
 * ncomp = 3 is the fused read of pm.py:54-56. */
int jpm_cic_read_f32(void *stream, int nx, int ny, int nz, const float *mesh,
                     int ncomp, int64_t comp_stride, const float *pos,
                     int64_t n, float *out, uint32_t flags);

/* Float64 CIC paint and read: jpm_cic_paint_f32 / jpm_cic_read_f32 in double
 * precision, the reference's computation under jax_enable_x64 (its
 * equivalence tests run in float64, tests/test_distributed_pm.py:30).
 * Deposits are summed with float64 atomics; JPM_DETERMINISTIC is rejected
 * (a float32 mode).  Flags: JPM_RELATIVE, JPM_ZERO_MESH.  Replaces
 * painting.py:15-45 / 78-106 and, relative, 161-189 / 218-236. */
int jpm_cic_paint_f64(void *stream, int nx, int ny, int nz, const double *pos,
                      int64_t n, const double *weight, double wscalar,
                      double *mesh, uint32_t flags);
int jpm_cic_read_f64(void *stream, int nx, int ny, int nz, const double *mesh,
                     int ncomp, int64_t comp_stride, const double *pos,
                     int64_t n, double *out, uint32_t flags);

/* Float64 pm_forces (pm.py:12-58 under jax_enable_x64): paint of the n
 * positions (or the real density `delta` when not NULL), hipFFT D2Z, the
 * force multipliers -g_d * invlaplace * longrange (k-space flags as for
 * jpm_kspace_force_f32) in double, three Z2D, read at the positions into
 * out3[n][3].  flags: JPM_RELATIVE for displacements.  Workspace:
 * jpm_pm_forces_f64_workspace_size bytes. */
size_t jpm_pm_forces_f64_workspace_size(int nx, int ny, int nz);
int jpm_pm_forces_f64(void *stream, int nx, int ny, int nz, const double *pos,
                      int64_t n, const double *delta, double r_split,
                      uint32_t flags, uint32_t kflags, double *out3,
                      void *ws, size_t ws_bytes);

/* Bytes of workspace jpm_cic_paint_multi_f32 needs (0 unless deterministic). */
size_t jpm_cic_paint_multi_workspace_size(int nx, int ny, int nz, int nc,
                                          uint32_t flags);

/* nc-channel paint (nc = 1..4): channel c deposits weights[p*nc + c] into
 * mesh + c*nx*ny*nz.  The transpose of the nc-component jpm_cic_read_f32;
 * it is the mesh cotangent of cic_read in the hand-written adjoint (the
 * reference gets it from XLA's transpose of the gather, painting.py:104-106). */
int jpm_cic_paint_multi_f32(void *stream, int nx, int ny, int nz,
                            const float *pos, int64_t n, const float *weights,
                            int nc, float *mesh, uint32_t flags, void *ws,
                            size_t ws_bytes);

/* Workspace bytes for jpm_cic_paint_binned_f32 (nc = 1 or 3). */
size_t jpm_cic_paint_binned_workspace_size(int nx, int ny, int nz, int64_t n,
                                           int nc, uint32_t flags);

/* The binned paint, the default inside every pipeline: a counting sort of
 * the particles into 16^3-cell tiles (LDS histograms, scan, LDS-ranked
 * scatter), one workgroup per chunk of a tile's particles accumulating into
 * the tile's 17^3 LDS window, and a gather-reduce of the windows into the
 * mesh.  No global float atomics; JPM_DETERMINISTIC makes the window sums
 * 64-bit fixed point (bitwise reproducible).  nc = 1 (weights per particle or
 * wscalar) or 3 (weights[p*3 + c], the adjoint of the 3-component read).
 * Replaces the scatter_add of painting.py:39-44 / painting_utils.py:110;
 * meshes of up to 32768 tiles count and scatter through per-workgroup LDS
 * histograms, larger ones through wavefront-aggregated global atomics. */
int jpm_cic_paint_binned_f32(void *stream, int nx, int ny, int nz,
                             const float *pos, int64_t n, const float *weights,
                             float wscalar, int nc, float *mesh,
                             uint32_t flags, void *ws, size_t ws_bytes);

/* Bytes of device scratch the allocation-free paint holds per GPU (64 MiB). */
#define JPM_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* jpm_cic_paint_binned_f32 without a caller workspace and without any device
 * allocation: the tile sort, windows and counters live in a file-scope device
 * array of JPM_NOALLOC_SCRATCH_BYTES per GPU, held from the library's load to
 * the end of the process (its address is looked up once per device, in
 * jpm_init or on the first call on that device).  Same arguments, flags,
 * results and bits as jpm_cic_paint_binned_f32.  A call whose workspace
 * (jpm_cic_paint_binned_workspace_size) exceeds the scratch returns
 * JPM_E_INVALID with the sizes in jpm_last_error(); 128^3 particles on a 128^3
 * mesh fit for nc = 1.  Every call on one GPU uses the same scratch, so calls
 * that overlap on one GPU (several streams or threads) race; calls ordered on
 * one stream do not.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc). */
int jpm_cic_paint_binned_noalloc_f32(void *stream, int nx, int ny, int nz,
                                     const float *pos, int64_t n,
                                     const float *weights, float wscalar,
                                     int nc, float *mesh, uint32_t flags);

/* Position cotangent of CIC sampling:
 *   gpos[p][e] (+)= w_p * sum_c cot[p*ncomp+c] * sum_o mesh_c[idx_o] dW_o/dx_e
 * with dW/dx from d|u|/du = +1 at u >= 0 (the jnp.abs convention).  cot NULL
 * means 1; weight NULL means the scalar wscalar; accumulate != 0 adds to gpos.
 * The position gradient of cic_read (painting.py:78-106) and cic_paint
 * (painting.py:15-45) under jax.grad. */
int jpm_cic_grad_pos_f32(void *stream, int nx, int ny, int nz,
                         const float *mesh, int ncomp, int64_t comp_stride,
                         const float *pos, int64_t n, const float *cot,
                         const float *weight, float wscalar, float *gpos,
                         uint32_t flags, int accumulate);

/* ------------------------------------------------------------------- FFT */

/* R2C of `batch` meshes (stride nx*ny*nz floats) into half spectra (stride
 * nx*ny*(nz/2+1) complex), unnormalised.  Replaces the forward c2c
 * jaxdecomp.pfft3d of distributed.py:37-38 for real input. */
int jpm_fft_r2c_f32(void *stream, int nx, int ny, int nz, int batch,
                    const float *in, void *out);

/* C2R of `batch` half spectra, unnormalised (the caller folds 1/M into the
 * k-space multiplier).  `in` is clobbered.  Replaces pifft3d(...).real of
 * distributed.py:41-42 for Hermitian spectra. */
int jpm_fft_c2r_f32(void *stream, int nx, int ny, int nz, int batch,
                    void *in, float *out);

/* Complex-to-complex transform of `batch` [nx][ny][nz] complex64 arrays,
 * direction -1 forward / +1 inverse, unnormalised; in == out allowed.
 * The c2c jaxdecomp.pfft3d / pifft3d of distributed.py:37-42 for complex
 * input.  Computed from one batched R2C of the real and imaginary parts
 * (Z = X + iY, the inverse as conj(FFT(conj Z))) in a workspace of
 * jpm_fft_c2c_workspace_size bytes: no complex-to-complex hipFFT plan is
 * ever created, because creating one changed the results of R2C plans of
 * other shapes created after it in the same process. */
size_t jpm_fft_c2c_workspace_size(int nx, int ny, int nz, int batch);
int jpm_fft_c2c_f32(void *stream, int nx, int ny, int nz, int batch,
                    const void *in, void *out, int direction, void *ws,
                    size_t ws_bytes);

/* ------------------------------------------------------------ k-space */

/* out3[d] = scale * (-g_d(k)) * G(k) * Lr(k) * [C(k)] * delta_k for d=0,1,2
 * (three consecutive half spectra).  G = invlaplace_kernel, g_d =
 * gradient_kernel, Lr = longrange_kernel(r_split), C = cic_compensation
 * (JPM_K_DECONV only).  Replaces pm.py:49-56 with kernels.py:41-115. */
int jpm_kspace_force_f32(void *stream, int nx, int ny, int nz,
                         const void *delta_k, void *out3, float r_split,
                         float scale, uint32_t flags);

/* out = scale * G(k) Lr(k) [C(k)] delta_k: the potential of pm.py:51-52
 * (one read, one write of the half spectrum; out may not alias delta_k).
 * With the default 4-point gradient the pipelines inverse-transform this
 * once and take the gradient in real space (jpm_fd4_gradient_f32). */
int jpm_kspace_potential_f32(void *stream, int nx, int ny, int nz,
                             const void *delta_k, void *out, float r_split,
                             float scale, uint32_t flags);

/* out3[d] = -scale * D_d phi with the periodic 4-point central difference
 * D f(c) = (8 (f(c+1) - f(c-1)) - (f(c+2) - f(c-2))) / 12, whose Fourier
 * multiplier is exactly gradient_kernel(order=1) = i (8 sin w - sin 2w) / 6
 * (kernels.py:63-66); -D_d IFFT[P] == IFFT[-g_d P] (pm.py:54-56). */
int jpm_fd4_gradient_f32(void *stream, int nx, int ny, int nz,
                         const float *phi, float scale, float *out3);

/* out = scale * sum_d D_d g3[d]: the transpose of jpm_fd4_gradient_f32
 * (D^T = -D), used by the force adjoint. */
int jpm_fd4_divergence_f32(void *stream, int nx, int ny, int nz,
                           const float *g3, float scale, float *out);

/* Multiply a half spectrum in place by scale * C(k) (cic_compensation,
 * kernels.py:118-136): the k-space step of compensate_cic,
 * painting.py:263-275. */
int jpm_kspace_compensate_f32(void *stream, int nx, int ny, int nz,
                              void *field_k, float scale);

/* field_k *= scale * sqrt(pk(|k| n/L) * M / V) with pk the linear
 * interpolation of the table (k_tab[ntab] ascending, pk_tab[ntab]) clamped at
 * the ends (jnp.interp).  k_tab/pk_tab are device arrays.  Replaces
 * pm.py:135-143 (linear_field). */
int jpm_kspace_linear_field_f32(void *stream, int nx, int ny, int nz,
                                void *field_k, float Lx, float Ly, float Lz,
                                const float *k_tab, const float *pk_tab,
                                int ntab, float scale);

/* Workspace bytes for jpm_kspace_linear_field_table_vjp_f32. */
size_t jpm_linear_field_table_vjp_workspace_size(int ntab);

/* Cotangent of pk_tab for jpm_kspace_linear_field_f32 followed by C2R: given
 * the noise half spectrum noise_k and the half spectrum cot_k of the real
 * output's cotangent (both unnormalised R2C), out[ntab] (device, float64) =
 * sum over the full spectrum of Re(W conj G) dc/dpk_tab, c the forward
 * multiplier.  The adjoint of pm.py:135-143 with respect to the pk table
 * (jax.grad of linear_field through jnp.interp).  Deterministic. */
int jpm_kspace_linear_field_table_vjp_f32(void *stream, int nx, int ny, int nz,
                                          const void *noise_k, const void *cot_k,
                                          float Lx, float Ly, float Lz,
                                          const float *k_tab, const float *pk_tab,
                                          int ntab, float scale, double *out,
                                          void *ws, size_t ws_bytes);

/* Six half spectra scale * g_i g_j G(k) delta_k for (i,j) in
 * (0,0),(1,1),(2,2),(0,1),(0,2),(1,2): the 2LPT shear terms of pm.py:89-109. */
int jpm_kspace_shear6_f32(void *stream, int nx, int ny, int nz,
                          const void *delta_k, void *out6, float scale,
                          uint32_t flags);

/* delta2 = sum_{i<j} (phi_ii phi_jj - phi_ij^2) from six meshes stored
 * `M` floats apart in the order of jpm_kspace_shear6_f32, accumulated in the
 * order of pm.py:92-109. */
int jpm_lpt2_source_f32(void *stream, int64_t M, const float *phi6,
                        float *delta2);

/* Adjoint of jpm_kspace_force_f32 as a real-linear map of half spectra:
 * out = scale * sum_d conj(-g_d G Lr [C]) in3[d].  Composed with R2C/C2R it
 * is the transpose of the force field operator of pm.py:49-56. */
int jpm_kspace_force_adjoint_f32(void *stream, int nx, int ny, int nz,
                                 const void *in3, void *out, float r_split,
                                 float scale, uint32_t flags);

/* Transpose of jpm_kspace_shear6_f32 (the multipliers are real and even):
 * out = scale * sum_m (-g_a g_b G) in6[m]. */
int jpm_kspace_shear6_adjoint_f32(void *stream, int nx, int ny, int nz,
                                  const void *in6, void *out, float scale,
                                  uint32_t flags);

/* Vector-Jacobian product of jpm_lpt2_source_f32: out6[m] = dL/dphi_m given
 * gbar = dL/d(delta2) (the adjoint of pm.py:92-109). */
int jpm_lpt2_source_vjp_f32(void *stream, int64_t M, const float *phi6,
                            const float *gbar, float *out6);

/* Hermitian part of a full c2c spectrum [nx][ny][nz] on the half grid,
 * (X(k) + conj(X(-k)))/2, so that C2R(half)/M == Re(ifftn(X)) for any X
 * (ifft3d of distributed.py:41-42; pm_forces with a complex delta,
 * pm.py:46-47). */
int jpm_spectrum_full_to_half_f32(void *stream, int nx, int ny, int nz,
                                  const void *full, void *half);

/* Full c2c spectrum of a real field from its half spectrum (fft3d of
 * distributed.py:37-38 for real input). */
int jpm_spectrum_half_to_full_f32(void *stream, int nx, int ny, int nz,
                                  const void *half, void *full);

/* ------------------------------------------------------------ analysis */

/* Workspace bytes for jpm_power_spectrum_f32 (per-workgroup partial sums). */
size_t jpm_power_spectrum_workspace_size(int nx, int ny, int nz, int nedges,
                                         int npoles, int cross);

/* Mode sums of the binned power spectrum, utils.py:76-128 (power_spectrum,
 * bins of _initialize_pk utils.py:14-73), over the R2C half spectra spec1 and
 * (cross spectrum, nullable) spec2 of real [nx][ny][nz] meshes, counting the
 * Hermitian mirror of every mode the half spectrum does not store.
 * |k| = float32 sqrt of the float32-rounded float64 sum of (fac_d *
 * fftfreq)^2, fac_d = 2 pi n_d / L_d; bin = np.digitize(|k|, edges) over the
 * nedges increasing float64 device edges (nedges + 1 bins).  out (device,
 * float64) is [slot][bin]: slot 0 mode count, 1 sum of |k|, 2+q the real
 * part of sum (2 l_q + 1) P_l_q(mu) X1 conj(X2) for pole poles[q] (host
 * array, npoles <= 8), and for cross spectra 2+npoles+q its imaginary part;
 * mu = k.los / |k| (0 at k = 0) when any pole is nonzero.  The FFTs are
 * unnormalised: divide the power sums by nx*ny*nz for norm='ortho'.
 * Deterministic. */
int jpm_power_spectrum_f32(void *stream, int nx, int ny, int nz,
                           const void *spec1, const void *spec2, double fac_x,
                           double fac_y, double fac_z, const double *edges,
                           int nedges, const int *poles, int npoles,
                           double los_x, double los_y, double los_z,
                           double *out, void *ws, size_t ws_bytes);

/* ------------------------------------------------------------- pipeline */

/* Workspace bytes for jpm_pm_forces_f32 on this mesh and particle count. */
size_t jpm_pm_forces_workspace_size(int nx, int ny, int nz, int64_t n,
                                    uint32_t flags);

/* forces[n][3] = pm_forces(pos): paint (or use delta_k if non-NULL), R2C,
 * jpm_kspace_force_f32, 3 C2R, 3-component read.  `scale` multiplies the
 * result (1 for pm_forces, 1.5*Omega_m for the ODE).  paint flags
 * (JPM_RELATIVE, JPM_DETERMINISTIC) in the low byte, k-space flags shifted
 * left by 8.  Replaces pm.py:12-58. */
int jpm_pm_forces_f32(void *stream, int nx, int ny, int nz, const float *pos,
                      int64_t n, const void *delta_k, float *forces,
                      float r_split, float scale, uint32_t flags, void *ws,
                      size_t ws_bytes);

/* Workspace bytes for jpm_force_field_f32. */
size_t jpm_force_field_workspace_size(int nx, int ny, int nz, int64_t n,
                                      uint32_t flags);

/* phi3[d] (three meshes, 3*nx*ny*nz floats) = scale * Re IFFT[-g_d G Lr FFT
 * rho], rho = paint(pos) (or delta_k when non-NULL): the force meshes of
 * pm.py:49-56 before the read.  Flags as jpm_pm_forces_f32. */
int jpm_force_field_f32(void *stream, int nx, int ny, int nz, const float *pos,
                        int64_t n, const void *delta_k, float *phi3,
                        float r_split, float scale, uint32_t flags, void *ws,
                        size_t ws_bytes);

/* Workspace bytes for jpm_force_vjp_f32. */
size_t jpm_force_vjp_workspace_size(int nx, int ny, int nz, int64_t n,
                                    uint32_t flags);

/* gpos[n][3] += scale * J(pos)^T cot, J the Jacobian of the pm_forces output
 * (pm.py:12-58) with respect to the particle positions, given the force
 * meshes phi3 of those positions (jpm_force_field_f32 with scale 1).  This is
 * the hand-written counterpart of jax.grad through pm_forces. */
int jpm_force_vjp_f32(void *stream, int nx, int ny, int nz, const float *phi3,
                      const float *pos, int64_t n, const float *cot,
                      float scale, float *gpos, float r_split, uint32_t flags,
                      void *ws, size_t ws_bytes);

/* cic_read (painting.py:78-106) through a tile sort of the positions, for
 * particles in no spatial order: absolute positions, 1 to 3 meshes
 * (comp_stride floats apart), out[p * ncomp + c], bit for bit the values of
 * jpm_cic_read_f32.  Workspace of jpm_cic_read_sorted_workspace_size bytes
 * (0: the mesh is too large for the tile sort). */
size_t jpm_cic_read_sorted_workspace_size(int nx, int ny, int nz, int64_t n);
int jpm_cic_read_sorted_f32(void *stream, int nx, int ny, int nz,
                            const float *mesh, int ncomp, int64_t comp_stride,
                            const float *pos, int64_t n, float *out,
                            uint32_t flags, void *ws, size_t ws_bytes);

/* Workspace bytes for jpm_force_kick_f32 (0 when three meshes exceed 2^32
 * cells or flags ask for the exact gradient / atomic paint, which the sorted
 * pipeline does not cover). */
size_t jpm_force_kick_workspace_size(int nx, int ny, int nz, int64_t n,
                                     uint32_t flags);

/* One force evaluation of the leapfrog on the tile-sorted particles:
 * phi (M floats, out) = Re IFFT[G Lr [C] FFT(paint(pos))] / M, then
 * vel[p] += gamma * read(-D phi, pos[p]) with the 4-point stencil D applied
 * in LDS (write_forces != 0: vel[p] = gamma * forces instead).  pm.py:12-58
 * fused with the kick of ode.py:142-143.  phi is what jpm_force_vjp_pot_f32
 * needs for the adjoint of this evaluation.  vel = NULL (with write_forces
 * = 0) stops after phi and the tile sort: the force evaluation alone, for an
 * adjoint that needs its data but not its kick.  counts_in (nullable):
 * per-tile particle counts of pos from the drift that wrote it
 * (jpm_force_kick_drift_f32's counts_out), which replace the sort's
 * counting pass (absolute positions only). */
int jpm_force_kick_f32(void *stream, int nx, int ny, int nz, const float *pos,
                       int64_t n, float *vel, float gamma, float *phi,
                       float r_split, uint32_t flags, int write_forces,
                       void *ws, size_t ws_bytes, void *keep_sort,
                       size_t keep_bytes, const uint32_t *counts_in);

/* Bytes of a tile sort of n particles under the pipeline `flags` (their
 * sort tuning bits): the `keep_sort` buffer of jpm_force_kick_f32 (optional;
 * when given the sort is built there and stays valid for
 * jpm_force_vjp_pot_f32 of the same positions and flags). */
size_t jpm_bin_sort_size(int nx, int ny, int nz, int64_t n, uint32_t flags);

/* Workspace bytes for jpm_force_vjp_pot_f32 (0 when unsupported, as above). */
size_t jpm_force_vjp_pot_workspace_size(int nx, int ny, int nz, int64_t n,
                                        uint32_t flags);

/* gpos[n][3] += scale * J(pos)^T cot for the forces of jpm_force_kick_f32,
 * given that evaluation's potential phi (and optionally its kept tile sort,
 * `sorted`, so the positions are not sorted again): the 3-channel paint of cot,
 * the stencil divergence, the potential operator and one LDS-window gather
 * for the direct and the density terms. */
int jpm_force_vjp_pot_f32(void *stream, int nx, int ny, int nz,
                          const float *phi, const float *pos, int64_t n,
                          const float *cot, float scale, float *gpos,
                          float r_split, uint32_t flags, void *ws,
                          size_t ws_bytes, const void *sorted,
                          size_t sorted_bytes);

/* The leapfrog's kick fused with the drift that follows it (ode.py:142-146:
 * a kick, then the next step's drift): jpm_force_kick_f32 with
 * write_forces = 0, then pos_out[p] = pos[p] + beta * vel[p] with the new
 * velocity, also written to pos_keep when it is not NULL.  pos_out may be
 * pos.  Same workspace and keep_sort as jpm_force_kick_f32.
 * Absolute positions: counts_out, when not NULL, receives the tile counts
 * of pos_out (jpm_tile_count(nx, ny, nz) uint32), and a later call whose
 * pos is that pos_out, unchanged, may pass them as counts_in, which skips
 * the counting pass of its tile sort (NULL: count).  The two arrays must
 * differ.  vel_out (nullable): the kicked velocities go there instead of
 * into vel, which is then only read (a checkpoint keeps them without a
 * copy); vel_out must not overlap vel. */
int jpm_force_kick_drift_f32(void *stream, int nx, int ny, int nz,
                             const float *pos, int64_t n, float *vel,
                             float gamma, float beta, float *pos_out,
                             float *pos_keep, float *phi, float r_split,
                             uint32_t flags, void *ws, size_t ws_bytes,
                             void *keep_sort, size_t keep_bytes,
                             const uint32_t *counts_in, uint32_t *counts_out,
                             float *vel_out);

/* Tiles (16^3 cells) of a mesh: the length of the tile-count arrays. */
size_t jpm_tile_count(int nx, int ny, int nz);

/* The adjoint of that fused step, reverse order: jpm_force_vjp_pot_f32 into
 * gpos, then the drift's adjoint cot[p] += beta * gpos[p] (cot updated in
 * place).  Same workspace as jpm_force_vjp_pot_f32.
 * cot_max_in: NULL, or a device float holding max_p,d |cot[p][d]| of the
 * cotangent given (the bound of the cotangent paint's fixed-point sums),
 * which skips that reduction; cot_max_out: NULL, or a device float that
 * receives max |cot| of the updated cotangent.  A leapfrog adjoint passes
 * the same float as both from its second step on (the value is exact, so
 * the results are the same bits as without it). */
int jpm_force_vjp_pot_drift_f32(void *stream, int nx, int ny, int nz,
                                const float *phi, const float *pos, int64_t n,
                                float *cot, float scale, float beta,
                                float *gpos, float r_split, uint32_t flags,
                                void *ws, size_t ws_bytes, const void *sorted,
                                size_t sorted_bytes, const float *cot_max_in,
                                float *cot_max_out);

/* Workspace bytes for jpm_dot_f32. */
size_t jpm_dot_workspace_size(void);

/* *out (device float64) = sum_i x[i] y[i], accumulated in float64 in a fixed
 * order: the scalar cotangents of the leapfrog coefficients and growth factors
 * (what jax.grad forms for the scalars of ode.py:139-143 and pm.py:84-124). */
int jpm_dot_f32(void *stream, int64_t n, const float *x, const float *y,
                double *out, void *ws, size_t ws_bytes);

/* y += a * x over n floats (16-byte aligned): the drift x += h v / (a^3 E)
 * of the leapfrog built on make_ode_fn (ode.py:139-140) and its adjoint. */
int jpm_axpy_f32(void *stream, int64_t n, float a, const float *x, float *y);

/* ------------------------------------------------------- lens planes */

/* mesh[nx][ny] += the weighted 2-D CIC deposits of n particles whose (x, y)
 * are pos[p * stride], pos[p * stride + 1] (stride >= 2, so the first two
 * coordinates of [n][3] positions can be painted in place): cic_paint_2d
 * (painting.py:131-158), corner order (0,0), (1,0), (0,1), (1,1), weights
 * (1 - |x - cx|)(1 - |y - cy|) times weight[p] when weight is not NULL,
 * cells int32(c) mod n.  Float atomics: the summation order varies, as
 * lax.scatter_add's does on a GPU. */
int jpm_cic_paint_2d_f32(void *stream, int nx, int ny, const float *pos,
                         int64_t n, int stride, const float *weight,
                         float *mesh);

/* Adjoint of jpm_cic_paint_2d_f32 for a mesh cotangent g[nx][ny] (what
 * jax.grad forms through painting.py:131-158): gpos (2 floats per particle
 * at stride gstride) = weight[p] sum_o g[idx_o] dW_o/d(x, y) and gweight[p] =
 * sum_o g[idx_o] W_o.  Either output may be NULL. */
int jpm_cic_paint_2d_vjp_f32(void *stream, int nx, int ny, const float *pos,
                             int64_t n, int stride, const float *weight,
                             const float *g, float *gpos, int gstride,
                             float *gweight);

/* density_plane without smoothing (lensing.py:11-40) into plane[res][res]:
 * (x, y) = jnp.mod(pos[:, :2], box_nx) / box_nx * res, weight 1 for
 * center - width/2 < z <= center + width/2 (the bounds rounded to float32),
 * 0 otherwise (those particles are skipped: their deposits are zeros), the
 * 2-D CIC paint into zeros, then division by (box_nx / res)(box_ny / res)
 * width. */
int jpm_density_plane_f32(void *stream, const float *pos, int64_t n,
                          int box_nx, int box_ny, double center, double width,
                          int res, float *plane);

/* Adjoint of jpm_density_plane_f32 for a plane cotangent g[res][res], which
 * is overwritten (divided by the normalisation): gpos[n][3] = the position
 * cotangent (z and the particles outside the slab get 0). */
int jpm_density_plane_vjp_f32(void *stream, const float *pos, int64_t n,
                              int box_nx, int box_ny, double center,
                              double width, int res, float *g, float *gpos);

/* Workspace bytes of jpm_gaussian_smoothing_f32 (one half spectrum). */
size_t jpm_gaussian_smoothing_workspace_size(int n0, int n1);

/* out[n0][n1] = Re ifft2(fft2(im) filter) with the filter of
 * utils.gaussian_smoothing (utils.py:208-222): exp(-|k|^2 / (2 s^2)),
 * |k| the norm of the fftfreq wave numbers, s = 1 / (2 pi sigma).  The
 * filter is real and even, so the operation is its own adjoint.  im and out
 * may be the same array. */
int jpm_gaussian_smoothing_f32(void *stream, int n0, int n1, const float *im,
                               float *out, double sigma, void *ws,
                               size_t ws_bytes);

/* Workspace bytes of jpm_tile_order_u32. */
size_t jpm_tile_order_workspace_size(int nx, int ny, int nz, int64_t n);

/* order[k] = index of the particle at slot k of the tile sort of absolute
 * positions pos[n][3]: particles grouped by 16^3-cell tile.  No counterpart
 * in the reference: the leapfrog driver (integrate.NBody, standing in for
 * the odeint/diffrax drive of make_ode_fn, ode.py:122-147, and its jax.grad)
 * runs its whole time loop on particles in this order, so that each tile's
 * particles are adjacent in memory, and returns results in the caller's
 * order.  The order within a tile may vary from call to call. */
int jpm_tile_order_u32(void *stream, int nx, int ny, int nz, const float *pos,
                       int64_t n, uint32_t *order, void *ws, size_t ws_bytes);

/* Rows of 3 floats reordered, out of place: out[k] = in[order[k]]
 * (scatter = 0) or out[order[k]] = in[k] (scatter = 1, the inverse), for
 * in0 -> out0 and, when in1 is not NULL, in1 -> out1 in the same launch.
 * Rows whose order entry is >= n are skipped.  No counterpart in the
 * reference (see jpm_tile_order_u32). */
int jpm_permute_rows_f32(void *stream, int64_t n, const uint32_t *order,
                         const float *in0, float *out0, const float *in1,
                         float *out1, int scatter);

/* vel[n][3] += gamma * read(phi3, pos): the kick of ode.py:142-143 fused
 * into the 3-component CIC read. */
int jpm_kick_f32(void *stream, int nx, int ny, int nz, const float *phi3,
                 const float *pos, int64_t n, float gamma, float *vel,
                 uint32_t flags);

/* -------------------------------------------------------------- batches
 *
 * Many simulations of different mesh sizes and cosmologies stepped together:
 * each stage of a force evaluation is one launch over the whole batch (the
 * FFTs and k-space passes one per mesh shape).  The reference runs
 * simulations one at a time (a Python loop over pm_forces / the integrator,
 * or jax.vmap over equal shapes); these entry points have no single
 * counterpart in it and serve the batched leapfrog of jaxpm_amd.integrate.
 * A batch concatenates its simulations: positions, velocities and
 * cotangents [N][3] (N = sum n_b), meshes [M] (M = sum nx ny nz), in the
 * order given; simulations of one mesh shape must be contiguous. */

/* Describe a batch: dims[b*3 .. b*3+2] the mesh of simulation b, counts[b]
 * its particles; flags JPM_RELATIVE for displacements (counts[b] = cells).
 * Writes an opaque handle to *batch (a device table is uploaded; this call
 * synchronises). */
int jpm_batch_create(int nsims, const int *dims, const int64_t *counts,
                     uint32_t flags, void **batch);
int jpm_batch_destroy(void *batch);

/* Tiles of all simulations of the batch (the tile-count arrays' length). */
size_t jpm_batch_tile_count(const void *batch);

/* Bytes of a kept tile sort / of the force_kick, force_vjp and dot
 * workspaces of this batch. */
size_t jpm_batch_sort_size(const void *batch);
size_t jpm_batch_force_kick_workspace_size(const void *batch);
size_t jpm_batch_force_vjp_workspace_size(const void *batch);
size_t jpm_batch_dot_workspace_size(const void *batch);

/* jpm_force_kick_f32 for every simulation at once: phi[M] receives each
 * simulation's potential, vel_b += gamma[b] F_b (or = gamma[b] F_b with
 * write_forces); gamma is a device array [nsims].  keep_sort (nullable)
 * receives the tile sort for jpm_batch_force_vjp_f32.  Replaces the
 * pm_forces (pm.py:12-58) + kick (ode.py:143) of each simulation.  vel =
 * NULL (with write_forces = 0) stops after phi and the sort, as in
 * jpm_force_kick_f32; gamma may then be NULL too.  counts_in (nullable): the
 * batch's per-tile counts of pos from the drift that wrote it. */
int jpm_batch_force_kick_f32(void *stream, const void *batch, const float *pos,
                             float *vel, const float *gamma, float *phi,
                             float r_split, uint32_t kflags, int write_forces,
                             void *ws, size_t ws_bytes, void *keep_sort,
                             size_t keep_bytes, const uint32_t *counts_in);

/* jpm_force_vjp_pot_f32 for every simulation: gpos_b += scale[b] J_b^T cot_b
 * given the evaluation's phi (and kept sort).  pot_scale[b] must be
 * float(scale[b] / (nx ny nz)_b), the factor the potential pass applies. */
int jpm_batch_force_vjp_f32(void *stream, const void *batch, const float *phi,
                            const float *pos, const float *cot,
                            const float *scale, const float *pot_scale,
                            float *gpos, float r_split, uint32_t kflags,
                            void *ws, size_t ws_bytes, const void *sorted,
                            size_t sorted_bytes);

/* The fused kick + drift and its adjoint (jpm_force_kick_drift_f32,
 * jpm_force_vjp_pot_drift_f32) for every simulation, beta a device array
 * [nsims]; counts_in / counts_out (jpm_batch_tile_count(batch) uint32),
 * vel_out and cot_max_in / cot_max_out as there, over the whole batch. */
int jpm_batch_force_kick_drift_f32(void *stream, const void *batch,
                                   const float *pos, float *vel,
                                   const float *gamma, const float *beta,
                                   float *pos_out, float *pos_keep, float *phi,
                                   float r_split, uint32_t kflags, void *ws,
                                   size_t ws_bytes, void *keep_sort,
                                   size_t keep_bytes,
                                   const uint32_t *counts_in,
                                   uint32_t *counts_out,
                                   float *vel_out);
int jpm_batch_force_vjp_drift_f32(void *stream, const void *batch,
                                  const float *phi, const float *pos,
                                  float *cot, const float *scale,
                                  const float *pot_scale, const float *beta,
                                  float *gpos, float r_split, uint32_t kflags,
                                  void *ws, size_t ws_bytes,
                                  const void *sorted, size_t sorted_bytes,
                                  const float *cot_max_in, float *cot_max_out);

/* y_b += a[b] x_b over each simulation's 3 n_b floats (the drift of
 * ode.py:140 with per-simulation coefficients). */
int jpm_batch_axpy_f32(void *stream, const void *batch, const float *a,
                       const float *x, float *y);

/* Workspace bytes of the batched initial-condition entry points below. */
size_t jpm_batch_lpt_workspace_size(const void *batch);

/* linear_field (pm.py:129-144) for every simulation: out_b = C2R[R2C(noise_b)
 * sqrt(pk_b(|k|) M_b / V_b)] / M_b with boxes[b*3..b*3+2] (device) and
 * pk_b = linear interpolation of (k_tab[ntab], pk_tabs[b*ntab..]) (device,
 * one k table for all).  Its own transpose: the noise cotangent is the same
 * map of the output cotangent. */
int jpm_batch_linear_field_f32(void *stream, const void *batch, const float *noise,
                               float *out, const float *boxes, const float *k_tab,
                               const float *pk_tabs, int ntab, void *ws,
                               size_t ws_bytes);

/* The first-order LPT force of lpt (pm.py:79-82) at the particle lattice:
 * psi[p][d] = Re IFFT[-g_d G FFT(field_b)] at lattice node p of simulation
 * b (one particle per cell in lattice order, the CIC read at a node being
 * the node value).  _vjp: gfield = the transpose applied to gpsi. */
int jpm_batch_lattice_forces_f32(void *stream, const void *batch, const float *field,
                                 float *psi, void *ws, size_t ws_bytes);
int jpm_batch_lattice_forces_vjp_f32(void *stream, const void *batch,
                                     const float *gpsi, float *gfield, void *ws,
                                     size_t ws_bytes);

/* jpm_tile_order_u32 over every simulation of an absolute-position batch
 * (workspace of jpm_batch_sort_size bytes); each simulation's slots stay
 * within its own particle range. */
int jpm_batch_tile_order_u32(void *stream, const void *batch, const float *pos,
                             uint32_t *order, void *ws, size_t ws_bytes);

/* cic_paint of every simulation's particles into zeroed meshes [M]
 * (workspace of jpm_batch_force_kick_workspace_size), and the position
 * cotangent gpos_p = sum_o gmesh[idx_o] dW_o/dx (painting.py:15-45 under
 * jax.grad). */
int jpm_batch_paint_f32(void *stream, const void *batch, const float *pos,
                        float *mesh, void *ws, size_t ws_bytes);
int jpm_batch_paint_vjp_f32(void *stream, const void *batch, const float *pos,
                            const float *gmesh, float *gpos);

/* out[b] (device float64) = <x_b, y_b> over each simulation's 3 n_b floats,
 * fixed-order float64 sums (the per-simulation coefficient cotangents). */
int jpm_batch_dot_f32(void *stream, const void *batch, const float *x,
                      const float *y, double *out, void *ws, size_t ws_bytes);

#ifdef __cplusplus
}
#endif
#endif /* JAXPM_HIP_H */
