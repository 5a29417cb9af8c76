/* C-ABI of the seistorch_amd HIP library (gfx950 / MI355X).
 *
 * Plain pointers and sizes only.  Every pointer argument is device memory
 * unless stated otherwise; `stream` is a hipStream_t passed as void*, and all
 * work is enqueued on it (nothing here synchronises the device).  Wavefields
 * are FP32 (B, nz, nx) with x contiguous, models and damping arrays (nz, nx),
 * exactly the tensors the reference hands to an equation backend.
 *
 * Every `int` function returns 0 on success and a nonzero code otherwise
 * (1 invalid argument, 2 HIP runtime error); seis_last_error() then returns a
 * message for the calling thread.  The Python layer turns a nonzero code into
 * RuntimeError, the reference's error type for bad tensors.
 */
#ifndef SEISTORCH_AMD_H
#define SEISTORCH_AMD_H

#include <stdint.h>

#define SEIS_ABI_VERSION 15

#ifdef __cplusplus
extern "C" {
#endif

/* Message for the last failing call on this thread. */
const char* seis_last_error(void);
int32_t seis_abi_version(void);

/* Launch timing (no reference counterpart; used by the benchmark).  When
 * enabled, the loop entry points record a HIP event pair on their stream
 * around every launch of the kernels below; seis_profile_collect waits for
 * the last event and returns the summed duration and the number of launches
 * of one kernel since the previous collect. */
#define SEIS_KERNEL_ACOUSTIC_FWD 0
#define SEIS_KERNEL_ACOUSTIC_BWD 1
#define SEIS_KERNEL_ELASTIC_FWD 2  /* stress + velocity stages of one step */
#define SEIS_KERNEL_ELASTIC_BWD 3  /* both adjoint stages of one step */
#define SEIS_KERNEL_HABC_FWD 4
#define SEIS_KERNEL_HABC_BWD 5
#define SEIS_KERNEL_BATCH_FWD 6  /* k_fwd_batch, direct launches only */
#define SEIS_KERNEL_BATCH_BWD 7  /* k_bwd_batch, direct launches only */
#define SEIS_KERNEL_ACOUSTIC_FWD2 8  /* k_fwd2: two forward steps per launch */
#define SEIS_KERNEL_ACOUSTIC_BWD2 9  /* k_bwd2: two reverse steps per launch */
#define SEIS_KERNEL_STREAM_COPY 10   /* k_stream_copy (seis_stream_copy) */
#define SEIS_PROFILE_KERNELS 11
int seis_profile_enable(int32_t on);
int seis_profile_collect(int32_t kernel, double* total_ms, int64_t* launches);

/* HBM calibration copy (no reference counterpart; used by the benchmark):
 * dst[i] = src[i] for n floats (n a multiple of 4, 16-byte aligned, no
 * overlap) with float4 accesses over a grid-stride loop of `blocks` blocks of
 * 256 threads.  Its achieved rate is the practical HBM ceiling, and its
 * FETCH_SIZE / WRITE_SIZE counters against the 4n bytes it moves each way
 * calibrate the counters of the other kernels. */
int seis_stream_copy(const float* src, float* dst, int64_t n, int32_t blocks, void* stream);

/* Number of cells in the boundary-saving ring of an (nz, nx) grid: rows
 * npml and nz-1-npml, columns npml and nx-1-npml, without the corners
 * (the 5-point neighbours of the undamped interior).  multiple != 0 is the
 * free surface: no top boundary, so the interior reaches row 0, the top row
 * npml is not part of the ring and the columns run from row 0 to nz-2-npml
 * (restore_boundaries(..., multiple=True) leaves the top row alone,
 * seistorch/equations2d/utils.py:71-78).  -1 if the grid has no interior.
 * Replaces the ring shape of seistorch/equations2d/utils.py:50-69. */
int64_t seis_ring_len(int64_t nz, int64_t nx, int32_t npml, int32_t multiple);

/* Number of field-pair snapshots the forward stores: floor((nt-2)/snap_every)
 * (0 if snap_every <= 0).  No reference counterpart: the reference keeps the
 * ring only (seistorch/checkpoint_new.py:137-139). */
int64_t seis_snapshot_count(int64_t nt, int32_t snap_every);

/* ---------------- acoustic, constant density, PML ------------------------ */

/* One time step y = A*(2/dt^2*h1 - B*h2 + vp^2*L h1), A = 1/(dt^-2 + d/dt),
 * B = dt^-2 - d/dt.  Replaces seistorch/equations2d/acoustic.py:73-86
 * (_time_step, whose conv2d Laplacian is acoustic.py:60-71).  y must not
 * alias h1. */
int seis_acoustic_step(const float* vp, const float* d, const float* h1, const float* h2,
                       float* y, int64_t B, int64_t nz, int64_t nx, float dt, float h,
                       void* stream);

/* Vector-Jacobian product of seis_acoustic_step for cotangent gy of y (and
 * gh1_pass of the pass-through output h1, may be NULL): writes gh1, gh2 and,
 * if gvp != NULL, the shot-summed gvp (nz, nx).  Replaces torch.autograd
 * through acoustic.py:81-84 (the reference's boundary_saving: false path,
 * seistorch/cell.py:74-75). */
int seis_acoustic_step_vjp(const float* vp, const float* d, const float* h1, const float* gy,
                           const float* gh1_pass, float* gh1, float* gh2, float* gvp, int64_t B,
                           int64_t nz, int64_t nx, float dt, float h, void* stream);

/* Whole forward loop (seistorch/rnn.py:178-212 with acoustic.py:73-86,
 * source.py:47-57 and probe.py:42-44): nt steps from zero fields, the
 * wavelet sample wavelet[n] (device, nt) added at src_zx[b] = (z, x) of each
 * shot after step n, receivers rec_flat[r] (flat indices into (B, nz, nx))
 * recorded into rec_out[n*R + r].  If ring != NULL the ring of u^{n+1} is
 * stored in slot n of ring ((nt-1) x B x seis_ring_len) for n < nt-1 (the
 * boundary saving of seistorch/checkpoint_new.py:114-143).  If snaps != NULL,
 * the pair (u^m, u^{m+1}) is also stored for every m = j*snap_every,
 * 1 <= j <= seis_snapshot_count(nt, snap_every), in slot j-1 of snaps
 * (count x 2 x B x nz x nx): the adjoint restarts its reverse-time
 * reconstruction from them, which bounds FP32 drift.  On return u_a holds
 * u^{nt} and u_b holds u^{nt-1}.  Launch shape: shots_per_group (1, 2, 4
 * or 8) shots share one thread (model reads amortised over them), each
 * thread owns rows_per_thread (1, 2 or 4) consecutive rows of vec_x (1, 2 or
 * 4; must divide nx, pointers vec_x*4-byte aligned) consecutive cells.
 * If u_c and u_d (two more (B, nz, nx) buffers, distinct from u_a, u_b) are
 * given, two time steps run per launch (the intermediate field stays in
 * registers), with the same results; NULL keeps one step per launch.
 * illum: NULL, or an (nz, nx) buffer that receives the source illumination
 * sum_n sum_b (u_b^{n+1})^2 over the post-source fields (rnn.py:127-128,
 * 204-205: zeroed per forward, the shots summed in order per time level and
 * then added), the divisor of seistorch/process.py:114-117.  illum_part:
 * NULL, or ceil(B/shots_per_group) more (nz, nx) planes with which the
 * two-step launches accumulate the illumination themselves (per shot group,
 * summed at the end: the same sum in another order) instead of a pass over
 * every shot's field per time level.  multiple != 0: free surface (the
 * reference's geom.multiple, seistorch/pml.py:49-53 damps no top strip), the
 * ring is that of seis_ring_len(nz, nx, npml, 1), without a top row.
 * extra_src / extra_wav / n_extra: n_extra more point sources (source
 * encoding, codingfwi.py:184-334 with source.py:54-55: many sources with
 * their own encoded wavelets in one field), rows (shot, z, x) of extra_src
 * (n_extra x 3 int32) with wavelets extra_wav (n_extra x nt); after step n,
 * extra_wav[t*nt + n] is added at each row's cell (cells shared by two
 * sources get both).  They need one step per launch (u_c = u_d = NULL);
 * n_extra = 0 (NULL tables) for none. */
int seis_acoustic_forward(const float* vp, const float* d, const float* wavelet, int64_t nt,
                          const int32_t* src_zx, const int32_t* extra_src,
                          const float* extra_wav, int64_t n_extra, const int64_t* rec_flat,
                          int64_t R, float* rec_out, float* ring, float* snaps, int32_t snap_every,
                          float* illum, float* illum_part, float* u_a, float* u_b, float* u_c,
                          float* u_d, int64_t B, int64_t nz, int64_t nx, float dt, float h,
                          int32_t npml, int32_t multiple, int32_t shots_per_group,
                          int32_t rows_per_thread, int32_t vec_x, void* stream);

/* Adjoint of seis_acoustic_forward (replaces the boundary-saving backward of
 * seistorch/checkpoint_new.py:146-217 and acoustic.py:88-137, with the exact
 * time pairing): rec_grad[n*R + r] = dJ/d rec_out[n*R + r]; ring and snaps
 * (may be NULL) as written by seis_acoustic_forward with the same
 * snap_every; u_a/u_b hold the final forward state (consumed); mu_a, mu_b are
 * (B, nz, nx) work buffers;
 * grad_part is ceil(B/shots_per_group) x (nz, nx) work; grad (nz, nx) gets
 * dJ/dvp on the interior [npml+1, n-npml-1) and zero elsewhere; with
 * multiple != 0 (free surface, as in the forward) the interior rows are
 * [0, nz-npml-1), the region that seistorch/equations2d/acoustic.py:114-137
 * rebuilds when it leaves the top ring row alone.  extra_src, extra_wav,
 * n_extra: the forward's extra sources (removed again by the reverse-time
 * rebuild; they must lie in the rebuilt interior); they turn off the
 * two-step adjoint.
 * If u_c, u_d, mu_c, mu_d (four more (B, nz, nx) buffers) and, when R > 0,
 * rec_map are given, two reverse steps run per launch where the ring is
 * available: rec_map (nz x nx int32) holds each cell's receiver index within
 * a shot or -1, which requires that every shot has its receivers at the same
 * cells in the same order (rec_flat[b*R/B + r] = b*nz*nx + cell_r) and no cell
 * twice.  Results equal the one-step path except for the order of the
 * gradient additions at receiver cells.  NULL keeps one step per launch. */
int seis_acoustic_backward(const float* vp, const float* d, const float* wavelet, int64_t nt,
                           const int32_t* src_zx, const int32_t* extra_src,
                           const float* extra_wav, int64_t n_extra, const int64_t* rec_flat,
                           int64_t R,
                           const float* rec_grad, const float* ring, const float* snaps,
                           int32_t snap_every, float* u_a, float* u_b,
                           float* mu_a, float* mu_b, float* u_c, float* u_d, float* mu_c,
                           float* mu_d, const int32_t* rec_map, float* grad_part, float* grad,
                           int64_t B, int64_t nz, int64_t nx, float dt, float h, int32_t npml,
                           int32_t multiple, int32_t shots_per_group, int32_t rows_per_thread,
                           int32_t vec_x, void* stream);

/* ---------------- acoustic PML, uneven multi-model batch ------------------ */
/* n independent acoustic PML problems advanced together: every time step is
 * one launch over the tiles of all instances still running (the reference
 * runs each model through its own rnn.py:178-212 loop and backward,
 * seistorch/checkpoint_new.py:146-217; these two functions are
 * seis_acoustic_forward / seis_acoustic_backward applied to every instance at
 * once, bit-identical per instance to the single-instance calls with the same
 * launch shape and snap_every).
 *
 * shape (host, n x 5 int64): nz, nx, B, nt, R of each instance;
 * dt_h (host, n x 2 float): dt, h.  Every device arena is the concatenation
 * over q = 0..n-1 (the caller's order) of the instance's array:
 *   vp, d, grad          (nz, nx)
 *   wavelet              (nt,)
 *   src_zx               (B, 2) int32
 *   rec_flat             (R,) int64, flat indices into the instance's (B, nz, nx)
 *   rec_out, rec_grad    (nt, R)
 *   ring                 (nt-1, B, seis_ring_len(nz, nx, npml, 0))
 *   snaps                (seis_snapshot_count(nt, snap_every), 2, B, nz, nx)
 *   u_a, u_b, mu_a, mu_b (B, nz, nx); u_c, u_d too, or both null
 *   grad_part            (ceil(B/shots_per_group), nz, nx)
 * table is device scratch of n x SEIS_BATCH_COLS int64 (written here).
 * The forward advances two time steps per launch when u_c and u_d are given
 * (four rotating field arenas; records, rings, snapshots and the final u_a,
 * u_b equal the one-step loop's bit for bit), one step per launch when both
 * are null.
 * vec_x must divide every nx.  use_graph != 0 records the whole time loop
 * into one hipGraph on an internal capture stream and launches it on
 * `stream` (waits for it to finish before returning); 0 launches the
 * kernels directly on `stream`. */
#define SEIS_BATCH_COLS 23
int seis_acoustic_forward_batch(int64_t n, const int64_t* shape, const float* dt_h,
                                const float* vp, const float* d, const float* wavelet,
                                const int32_t* src_zx, const int64_t* rec_flat, float* rec_out,
                                float* ring, float* snaps, int32_t snap_every, float* u_a,
                                float* u_b, float* u_c, float* u_d, int64_t* table, int32_t npml,
                                int32_t shots_per_group, int32_t rows_per_thread, int32_t vec_x,
                                int32_t use_graph, void* stream);

int seis_acoustic_backward_batch(int64_t n, const int64_t* shape, const float* dt_h,
                                 const float* vp, const float* d, const float* wavelet,
                                 const int32_t* src_zx, const int64_t* rec_flat,
                                 const float* rec_grad, const float* ring, const float* snaps,
                                 int32_t snap_every, float* u_a, float* u_b, float* mu_a,
                                 float* mu_b, float* grad_part, float* grad, int64_t* table,
                                 int32_t npml, int32_t shots_per_group, int32_t rows_per_thread,
                                 int32_t vec_x, int32_t use_graph, void* stream);

/* ---------------- acoustic, constant density, HABC ----------------------- */
/* b is the (nz, nx) HABC weight field (seistorch/habc.py:42-81); the strips
 * are 50 cells wide; multiple != 0 selects the free-surface variant (no top
 * strip).  Grids need at least 102 cells per axis. */

/* One HABC step y = HABC(h1, h2) (acoustic_habc.py:206-221 with habc()
 * :147-204, including the torch.roll wrap of the 51-row strips).  y must not
 * alias h1 or h2. */
int seis_habc_step(const float* vp, const float* b, const float* h1, const float* h2, float* y,
                   int64_t B, int64_t nz, int64_t nx, float dt, float h, int32_t multiple,
                   void* stream);

/* VJP of seis_habc_step: gh1 = (dy/dh1)^T gy, gh2 = (dy/dh2)^T gy and, if gvp
 * != NULL, the shot-summed (dy/dvp)^T gy over every cell (replaces autograd
 * through acoustic_habc.py:206-221). */
int seis_habc_step_vjp(const float* vp, const float* b, const float* h1, const float* h2,
                       const float* gy, float* gh1, float* gh2, float* gvp, int64_t B, int64_t nz,
                       int64_t nx, float dt, float h, int32_t multiple, void* stream);

/* Whole HABC forward loop, as seis_acoustic_forward (same records, ring,
 * snapshot and illum layouts) with three field buffers: on return u_a holds
 * u^{nt}, u_b u^{nt-1}, u_c is scratch. */
int seis_habc_forward(const float* vp, const float* b, const float* wavelet, int64_t nt,
                      const int32_t* src_zx, const int64_t* rec_flat, int64_t R, float* rec_out,
                      float* ring, float* snaps, int32_t snap_every, float* illum, float* u_a,
                      float* u_b, float* u_c, int64_t B, int64_t nz, int64_t nx, float dt,
                      float h, int32_t npml, int32_t multiple, int32_t shots_per_group,
                      void* stream);

/* Adjoint of seis_habc_forward with boundary-saving reconstruction
 * (acoustic_habc.py:223-268 + checkpoint_new.py, exact pairing): three
 * cotangent buffers l_a, l_b, l_c; grad gets dJ/dvp on the interior. */
int seis_habc_backward(const float* vp, const float* b, const float* wavelet, int64_t nt,
                       const int32_t* src_zx, const int64_t* rec_flat, int64_t R,
                       const float* rec_grad, const float* ring, const float* snaps,
                       int32_t snap_every, float* u_a, float* u_b, float* l_a, float* l_b,
                       float* l_c, float* grad_part, float* grad, int64_t B, int64_t nz,
                       int64_t nx, float dt, float h, int32_t npml, int32_t multiple,
                       int32_t shots_per_group, void* stream);

/* ---------------- variable-density acoustic equation (HABC) --------------- */
/* One step of seistorch/equations2d/acoustic_rho_habc.py:31-57 (blend = 1:
 * _time_step, the HABC strips of b applied) or of its reverse step :59-89
 * before the ring restore and the source (blend = 0, b may be NULL):
 * y = 2 h1 - h2 + dt^2 h^-2 rho vp^2 (sum over x and z of
 * b_pos (h1[+1] - h1) - b_neg (h1 - h1[-1])), b = 1/rho averaged over the
 * two cells of each face, zero outside the grid.  vp, rho, b: (nz, nx);
 * h1, h2, y: (B, nz, nx).  dt and h are doubles: the constant dt^2 h^-2 is
 * formed in double, as from the reference's configuration scalars. */
int seis_rho_habc_step(const float* vp, const float* rho, const float* b, const float* h1,
                       const float* h2, float* y, int64_t B, int64_t nz, int64_t nx, double dt,
                       double h, int32_t multiple, int32_t blend, void* stream);

/* Whole forward loop of the variable-density equation (rnn.py:178-212 with
 * acoustic_rho_habc.py:31-57): nt steps from zero fields, wavelet[n] added at
 * src_zx[b] after step n, receivers rec_flat[r] recorded into
 * rec_out[n*R + r]; ring and snaps (may be NULL) as in seis_habc_forward
 * (ring: (nt-1) x B x seis_ring_len(nz, nx, npml, multiple), snaps:
 * count x 2 x B x nz x nx); buoy: (nz, nx) work buffer (receives 1/rho,
 * formed once per call); illum: NULL or the source illumination as in
 * seis_acoustic_forward; u_a, u_b, u_c: (B, nz, nx) rotating buffers
 * (u^{nt} ends in u_a, u^{nt-1} in u_b); shots_per_group >= 1 shots share
 * one thread (the per-cell set-up runs once for them). */
int seis_rho_habc_forward(const float* vp, const float* rho, const float* b,
                          const float* wavelet, int64_t nt, const int32_t* src_zx,
                          const int64_t* rec_flat, int64_t R, float* rec_out, float* ring,
                          float* snaps, int32_t snap_every, float* buoy, float* illum, float* u_a,
                          float* u_b, float* u_c, int64_t B, int64_t nz, int64_t nx, double dt,
                          double h, int32_t npml, int32_t multiple, int32_t shots_per_group,
                          void* stream);

/* Adjoint of seis_rho_habc_forward with boundary-saving reconstruction
 * (replaces the reference's checkpointed backward of
 * acoustic_rho_habc.py:31-121 through seistorch/checkpoint_new.py:146-217,
 * with the exact time pairing): rec_grad[n*R + r] = dJ/d rec_out[n*R + r];
 * ring and snaps as written by the forward with the same snap_every; u_a,
 * u_b: the forward's final state (consumed); l_a, l_b, l_c: (B, nz, nx)
 * cotangent buffers; buoy: (nz, nx) work; grad_part: 2 x ceil(B/S) x (nz, nx)
 * work; grad_vp and grad_rho (nz, nx) get dJ/dvp and dJ/drho on the undamped
 * interior (rows from 0 with multiple != 0) and zero elsewhere. */
int seis_rho_habc_backward(const float* vp, const float* rho, const float* b,
                           const float* wavelet, int64_t nt, const int32_t* src_zx,
                           const int64_t* rec_flat, int64_t R, const float* rec_grad,
                           const float* ring, const float* snaps, int32_t snap_every, float* buoy,
                           float* u_a, float* u_b, float* l_a, float* l_b, float* l_c,
                           float* grad_part, float* grad_vp, float* grad_rho, int64_t B,
                           int64_t nz, int64_t nx, double dt, double h, int32_t npml,
                           int32_t multiple, int32_t shots_per_group, void* stream);

/* ---------------- elastic, isotropic velocity-stress, PML ---------------- */
/* Fields are passed as an array of five (B, nz, nx) pointers in the order
 * vx, vz, txx, tzz, txz (seistorch/eqconfigure.py Wavefield.elastic).  The
 * model enters through six per-cell coefficient planes written by
 * seis_elastic_coefficients: lam+2mu, lam, dt*mu/h, dt/(rho h), 1/(1+c), 1-c
 * with lam = rho(vp^2 - 2vs^2), mu = rho vs^2, c = dt*d/2 (elastic.py:11-14). */

/* coef (6 x nz x nx) from vp, vs, rho, d with the reference's FP32 order. */
int seis_elastic_coefficients(const float* vp, const float* vs, const float* rho, const float* d,
                              float* coef, int64_t nz, int64_t nx, float dt, float h,
                              void* stream);

/* One stage of the step, in place on the cells z0 <= z < z1, x0 <= x < x1:
 * stage 0 updates the stresses from the velocities (elastic.py:16-24),
 * stage 1 the velocities from the stresses (elastic.py:26-35); sign +1 is the
 * forward step (stage 0 then 1, whole grid), sign -1 on the interior region
 * is the reference's reverse step (elastic.py:39-122: stage 1 then 0). */
int seis_elastic_stage(int32_t stage, const float* coef, float* const* fields, float sign,
                       int32_t z0, int32_t z1, int32_t x0, int32_t x1, int64_t B, int64_t nz,
                       int64_t nx, float dt, float h, int32_t shots_per_group, void* stream);

/* Vector-Jacobian product of seis_elastic_stage (replaces torch.autograd
 * through elastic.py:7-37): cot_in are cotangents of the stage's outputs,
 * cot_out receives those of its inputs (no aliasing), fwd holds the stage's
 * input fields.  If grad_coef != NULL, the gradient w.r.t. the coefficient
 * planes is accumulated into it, grouped by shots_per_group: block q (4
 * planes: lam+2mu, lam, dt*mu/h, dt/(rho h)) for shots q*S .. q*S+S-1. */
int seis_elastic_stage_vjp(int32_t stage, const float* coef, float* const* cot_in,
                           float* const* cot_out, float* const* fwd, float* grad_coef,
                           float sign, int32_t z0, int32_t z1, int32_t x0, int32_t x1, int64_t B,
                           int64_t nz, int64_t nx, float dt, float h, int32_t shots_per_group,
                           void* stream);

/* Sum the `groups` coefficient-gradient blocks and chain them to dJ/dvp,
 * dJ/dvs, dJ/drho (nz x nx each). */
int seis_elastic_chain(const float* grad_coef, int32_t groups, const float* vp, const float* vs,
                       const float* rho, float* gvp, float* gvs, float* grho, int64_t nz,
                       int64_t nx, float dt, float h, void* stream);

/* Whole forward loop (rnn.py:176-212 with elastic.py:7-37, time offset 0):
 * per step the two stages, then wavelet[n] added to every field whose bit is
 * set in src_mask (bit f = field f) at src_zx[b] (source.py:47-52), then
 * channel c of receiver r recorded from field rec_field[c] into
 * rec_out[(n*R + r)*nch + c] (probe.py:42-44).  ring (if not NULL): the rings
 * of all five fields of state n+1 in slot n ((nt-1) x 5 x B x ring_len);
 * snaps (if not NULL): state m for m = j*snap_every <= nt-1 in slot j-1
 * (5 x B x nz x nx each).  fields are zeroed first and hold state nt on
 * return.  If scratch (five more (B, nz, nx) buffers) is not NULL, each step
 * runs as one fused pass (both stages, ping-ponging between fields and
 * scratch) with the same results.  illum (if not NULL): (nz, nx), receives
 * the source illumination sum_n sum_b of the squared post-source field
 * illum_field (0..4 = vx, vz, txx, tzz, txz; the reference uses its last
 * source type, rnn.py:204-205). */
int seis_elastic_forward(const float* coef, const float* wavelet, int64_t nt,
                         const int32_t* src_zx, int32_t src_mask, const int64_t* rec_flat,
                         int64_t R, const int32_t* rec_field, int32_t nch, float* rec_out,
                         float* ring, float* snaps, int32_t snap_every, float* illum,
                         int32_t illum_field, float* const* fields, float* const* scratch,
                         int64_t B, int64_t nz, int64_t nx, float dt, float h, int32_t npml,
                         int32_t multiple, int32_t shots_per_group, void* stream);

/* Adjoint of seis_elastic_forward with boundary-saving reconstruction
 * (replaces seistorch/checkpoint.py:108-230 and elastic.py:39-122 with the
 * exact time pairing).  fields: the final state from the forward (consumed);
 * adj_a, adj_b: two sets of five (B, nz, nx) work buffers; scratch: NULL,
 * or five more (B, nz, nx) buffers, with which each reverse step runs as one
 * fused pass (both transposed stages and the rebuild of the forward state,
 * ping-ponging between fields and scratch) with the same results; grad_coef:
 * ceil(B/S) x 4 planes, receives the coefficient gradients on the interior
 * (pass them to seis_elastic_chain). */
int seis_elastic_backward(const float* coef, const float* wavelet, int64_t nt,
                          const int32_t* src_zx, int32_t src_mask, const int64_t* rec_flat,
                          int64_t R, const int32_t* rec_field, int32_t nch,
                          const float* rec_grad, const float* ring, const float* snaps,
                          int32_t snap_every, float* const* fields, float* const* adj_a,
                          float* const* adj_b, float* const* scratch, float* grad_coef, int64_t B,
                          int64_t nz, int64_t nx, float dt, float h, int32_t npml,
                          int32_t multiple, int32_t shots_per_group, void* stream);

/* ---------------- multiscale record filtering ----------------------------- */
/* Zero-phase IIR filter of traces along time: replaces seistorch/signal.py:91-96
 * (backend 'torch', called from seistorch_dist.py:222-225), i.e.
 * torchaudio.functional.filtfilt(x.double(), a, b, clamp=False).float(): an
 * lfilter with zero initial state over time, then over reversed time, in
 * FP64, without edge padding.  x, y: (nt, ntr) float32, time-major, traces
 * contiguous (the records layout; y may be x); work: (nt, ntr) FP64
 * scratch; b, a: host arrays of ncoef FP64 coefficients (scipy.signal.butter
 * output, a[0] != 0), 1 <= ncoef <= 32.  Self-adjoint: the gradient of a
 * loss through the filter is the same call on the cotangent. */
int seis_filtfilt(const float* x, float* y, double* work, int64_t nt, int64_t ntr,
                  const double* b, const double* a, int32_t ncoef, void* stream);

/* ---------------- gradient post-processing -------------------------------- */
/* One 1-D Gaussian pass over a 2-D field: replaces seistorch/signal.py:247-304
 * (gaussian_filter, 2-D case).  x: (nz, nx) float32; w: host array of ks
 * FP32 taps (the caller builds them as the reference does: ks = 2r+1 for an
 * even radius r, 2r for an odd one, exp(-(k - ks/2)^2 / (2 sigma^2))
 * normalised in FP32); axis 0 = z, 1 = x.  y receives
 * y[i] = sum_k w[k] x[refl(i + k - ks/2)] along the axis (numpy 'reflect'),
 * and has n + 2*(ks/2) - ks + 1 cells along it: one more than x when ks is
 * even, as in the reference.  Requires ks/2 < the axis length; y must not
 * alias x. */
int seis_gaussian_filter(const float* x, float* y, int64_t nz, int64_t nx, int32_t axis,
                         const float* w, int32_t ks, void* stream);

/* The post-processing between the adjoint and the optimizer step, in the
 * order of seistorch_dist.py:270-280: replaces seistorch/process.py:66-112
 * (smooth_gradient, whose seabed pre-mask is pre_mask), :43-47
 * (cut_gradient: post_mask) and :114-117 (precondition: illum).  Each step
 * runs when its argument is given (non-NULL, counts > 0):
 *   g = grad * pre_mask;
 *   counts times: g = z pass (kz taps wz), then x pass (kx taps wx);
 *   g = g * post_mask;
 *   g = g / illum.
 * grad, out, work, pre_mask, post_mask, illum: (nz, nx) float32 (work unused
 * when counts == 0); out and work must not alias grad or each other.  kz and
 * kx must be odd (an even size, i.e. an odd radius, lengthens the gradient by
 * one cell per pass, which the reference's own callers cannot use) and
 * kz/2 < nz, kx/2 < nx. */
int seis_smooth_gradient(const float* grad, float* out, float* work, int64_t nz, int64_t nx,
                         const float* wz, int32_t kz, const float* wx, int32_t kx, int32_t counts,
                         const float* pre_mask, const float* post_mask, const float* illum,
                         void* stream);

/* seis_smooth_gradient without the work argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * work plane is a static workspace of SEIS_NOALLOC_MAX_CELLS FP32 cells
 * (16 MiB) inside the library, one per device, held until the process ends;
 * nothing is allocated and results equal seis_smooth_gradient's bit for bit.
 * Grids of more than SEIS_NOALLOC_MAX_CELLS cells are rejected (code 1).
 * Calls that overlap on one device share the workspace and race: serialise
 * them (one stream, or events). */
#define SEIS_NOALLOC_MAX_CELLS (1LL << 22)
int seis_smooth_gradient_noalloc(const float* grad, float* out, int64_t nz, int64_t nx,
                                 const float* wz, int32_t kz, const float* wx, int32_t kx,
                                 int32_t counts, const float* pre_mask, const float* post_mask,
                                 const float* illum, void* stream);

/* ---------------- learned gradient preconditioner (U-Net) ----------------- */
/* No reference counterpart: a small U-Net maps (gradient, model) to an
 * update direction inside FWI; the reference's plain-gradient step
 * (seistorch/fwi.py) is its fallback.  Activations are NHWC; bf16 tensors are
 * raw uint16 bit patterns (16-byte aligned); accumulation is FP32 on the
 * matrix cores. */

/* 3x3 convolution, zero padding 1: x (N,H,W,Cin) bf16, Cin a multiple of 8;
 * w (Cout, 9*Cin) bf16 with k = tap*Cin + ci, tap = 3*(dy+1) + (dx+1);
 * bias (Cout) FP32 or NULL.  Either y (bf16, = act(acc + bias), act 1 =
 * GELU(erf), optional z = acc + bias in bf16) or y_f32 (= acc + bias). */
int seis_unet_conv3x3(const uint16_t* x, int32_t N, int32_t H, int32_t W, int32_t Cin,
                      const uint16_t* w, const float* bias, int32_t Cout, int32_t act,
                      uint16_t* z, uint16_t* y, float* y_f32, void* stream);
/* wt (Cin, 9*Cout) = w with taps flipped and channel roles swapped: the data
 * gradient of a convolution is seis_unet_conv3x3 of its output gradient with wt. */
int seis_unet_flip_weights(const uint16_t* w, int32_t Cout, int32_t Cin, uint16_t* wt,
                           void* stream);
/* dw (Cout, 9*Cin) += weight gradient; db (Cout) += bias gradient (may be NULL). */
int seis_unet_conv3x3_wgrad(const uint16_t* x, const uint16_t* dz, int32_t N, int32_t H,
                            int32_t W, int32_t Cin, int32_t Cout, float* dw, float* db,
                            void* stream);
/* dz = bf16(dy * GELU'(z)). */
int seis_unet_gelu_bwd(const float* dy, const uint16_t* z, int64_t n, uint16_t* dz, void* stream);
int seis_unet_to_bf16(const float* x, int64_t n, uint16_t* y, void* stream);
/* 2x2 average pooling and its gradient (dx = or += 0.25*dy[pixel/2]). */
int seis_unet_avgpool2(const uint16_t* x, int32_t N, int32_t H, int32_t W, int32_t C, uint16_t* y,
                       void* stream);
int seis_unet_avgpool2_bwd(const float* dy, int32_t N, int32_t H, int32_t W, int32_t C, float* dx,
                           int32_t accumulate, void* stream);
/* out (N,H,W,C1+C2) = [nearest 2x upsampling of low (N,H/2,W/2,C1), skip (N,H,W,C2)]
 * and its gradient (dlow =, dskip = or +=). */
int seis_unet_up_concat(const uint16_t* low, const uint16_t* skip, int32_t N, int32_t H,
                        int32_t W, int32_t C1, int32_t C2, uint16_t* out, void* stream);
int seis_unet_up_concat_bwd(const float* dout, int32_t N, int32_t H, int32_t W, int32_t C1,
                            int32_t C2, float* dlow, float* dskip, int32_t accumulate_skip,
                            void* stream);
/* Network input (N,H,W,8) bf16: channel 0 = grad*scales[3n], channel 1 =
 * (model - scales[3n+2])*scales[3n+1], channels 2..7 zero. */
int seis_unet_pack_input(const float* grad, const float* model, int32_t N, int32_t H, int32_t W,
                         const float* scales, uint16_t* out, void* stream);
/* 1x1 head to one FP32 channel and its gradient (dw, db accumulate). */
int seis_unet_head(const uint16_t* x, int64_t P, int32_t C, const float* w, const float* b,
                   float* out, void* stream);
int seis_unet_head_bwd(const uint16_t* x, const float* dout, int64_t P, int32_t C, const float* w,
                       float* dx, float* dw, float* db, void* stream);
/* loss[0] = sum(mask*(pred-target))^2 * inv_count; dpred = its gradient. */
int seis_unet_mse(const float* pred, const float* target, const float* mask, int64_t n,
                  float inv_count, float* loss, float* dpred, void* stream);
/* AdamW step `step` (>= 1) on FP32 master weights, refreshing param_bf16 if not NULL. */
int seis_unet_adamw(float* param, const float* grad, float* m, float* v, int64_t n, float lr,
                    float beta1, float beta2, float eps, float weight_decay, int32_t step,
                    uint16_t* param_bf16, void* stream);
int seis_unet_fill(float* p, int64_t n, float v, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SEISTORCH_AMD_H */
