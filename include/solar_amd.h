/*
 * solar_amd.h -- C ABI of the MI355X (gfx950) acoustic RK4 hot path.
 *
 * Replaces the reference's `call RK4` (cube.f90:74) and everything below it
 * (timestep.f90:41-86): COMPUTE_RHS (compute_rhs_acoustic.f90:1), BC_RHS
 * (bc_acoustic_wave.f90:1), SPONGE (sponge.f90:1) and FILTER_EXP
 * (filter.f90:9), with the MPI panel exchange done on the device.
 *
 * Conventions
 *   - State q is the reference's q(mx,my,mr,5) for every panel, stacked with
 *     the panel slowest: q(n,n,nr,5,6), column-major, i (xi) fastest.  Fields
 *     1..5 = rho', v_xi, v_eta, v_r, p' (compute_rhs_acoustic.f90:20-25).
 *     Every panel block is one reference rank's block at px=py=pr=1.
 *   - Host pointers are caller-owned; the library copies them.  `_device`
 *     entry points take device pointers on the library's HIP device.
 *   - Every call returns 0 on success and a negative code on failure;
 *     solar_last_error() describes the last failure of the calling thread.
 *     (The reference returns no status: it STOPs, init_mpi.f90:31.)
 *   - A handle is used by one host thread at a time; all work is ordered on
 *     the handle's HIP stream.
 */
#ifndef SOLAR_AMD_H
#define SOLAR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct solar_ctx solar_ctx;

enum {
  SOLAR_OK = 0,
  SOLAR_ERR_ARG = -1,      /* invalid argument or configuration */
  SOLAR_ERR_HIP = -2,      /* HIP runtime error (no device, out of memory, launch) */
  SOLAR_ERR_NONFINITE = -3 /* non-finite value found by solar_check_finite */
};

/* Flags of solar_compute_rhs. */
#define SOLAR_RHS_BC 1     /* apply BC_RHS (bc_acoustic_wave.f90:14-34) */
#define SOLAR_RHS_SPONGE 2 /* apply SPONGE (sponge.f90:20-50) */

/*
 * Set-up of the time-invariant tables and device buffers.  Replaces
 * SETUP_COORDINATES_METRICS_ALL (metric.f90:8), SETUP_LAGRANGE_INTERPOLANT
 * (interpolation.f90:9), the partner-rank tables of INITIALIZE_MPI
 * (init_mpi.f90:164-304) and the background arrays SOLAR_DATA fills
 * (ic_acoustic_wave.f90:94-108, radial-only there, so passed as profiles).
 *   n          angular points per panel edge (nx = ny), 16 <= n <= 512 (any n;
 *              the sensitivity entry points and ensembles need a multiple of 16)
 *   nr         radial points, >= 12, or 1: the reference's 2-D case (no radial
 *              derivative, FILTER_R, BC_RHS or SPONGE, and r = 1: derivative.f90:620-623,
 *              filter.f90:647-650, bc_acoustic_wave.f90:11, sponge.f90:18,
 *              metric.f90:35-38; forward only)
 *   gr, c2, rho0, g0   radial profiles, length nr (nondimensional)
 *   sponge_len, a_sponge   sponge.f90:27,20 (k_sponge_length, a_sponge)
 */
int solar_init(solar_ctx** ctx, int n, int nr, const double* gr, const double* c2, const double* rho0,
               const double* g0, int sponge_len, double a_sponge);

/* As solar_init, choosing the storage type of the state and of the 2N
 * accumulator: state_bits = 64 (FP64, the reference's real*8) or 32 (FP32
 * storage; all arithmetic, metric terms, interpolation weights and ghost
 * strips stay FP64).  No reference counterpart: the reference is FP64 only. */
int solar_init_ex(solar_ctx** ctx, int n, int nr, const double* gr, const double* c2, const double* rho0,
                  const double* g0, int sponge_len, double a_sponge, int state_bits);

/* Storage width of the handle's state (64 or 32). */
int solar_state_bits(solar_ctx* ctx, int* bits);

/* Releases the handle and its device memory (MPI_FINALIZE, cube.f90:94). */
int solar_finalize(solar_ctx* ctx);

/* Copies the host state q(n,n,nr,5,6) to the device state (the COMMON
 * /RKWORK/ q that GENERATE_IC fills, cube.f90:47). */
int solar_set_state(solar_ctx* ctx, const double* q);

/* Copies the device state back to the host. */
int solar_get_state(solar_ctx* ctx, double* q);

/* Device-pointer state transfer (device-to-device copies on the handle's
 * stream); the device array holds q(n,n,nr,5,6) in the handle's storage type
 * (double, or float for a 32-bit handle). */
int solar_set_state_device(solar_ctx* ctx, const void* q_dev);
int solar_get_state_device(solar_ctx* ctx, void* q_dev);

/* nsteps calls of RK4 (timestep.f90:41-86): five stages of
 * COMPUTE_RHS -> BC_RHS -> SPONGE -> 2N update, then FILTER_EXP, then
 * time += dt.  *time (may be NULL) is advanced by nsteps*dt on the host.
 * Synchronous: returns when the device state is final. */
int solar_rk4(solar_ctx* ctx, int nsteps, double dt, double* time);

/* Enqueues nsteps RK4 steps on the handle's stream without waiting (the
 * timed loop of a benchmark); solar_sync waits for it. */
int solar_rk4_async(solar_ctx* ctx, int nsteps, double dt);
int solar_sync(solar_ctx* ctx);

/* One right-hand side: COMPUTE_RHS(q, q_rhs) (compute_rhs_acoustic.f90:1-159,
 * forcing term omitted), then BC_RHS and SPONGE as selected by flags.  Host
 * arrays q(n,n,nr,5,6) -> q_rhs(n,n,nr,5,6).  Does not touch the state. */
int solar_compute_rhs(solar_ctx* ctx, const double* q, double* q_rhs, int flags);

/* solar_compute_rhs on device arrays, allocating nothing: no hipMalloc,
 * hipFree or other allocating call.  No counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  q and q_rhs are device arrays
 * q(n,n,nr,5,6) in the handle's storage type (q_rhs must not overlap q);
 * flags as solar_compute_rhs.  The ghost strips of the panel exchange live in
 * a static device array of the library (128 MiB, held by every GPU the
 * library runs on until the process ends), so n*nr may be at most 116508
 * (6x256x256x384 fits); larger grids return SOLAR_ERR_ARG.  Enqueued on
 * `stream` (a hipStream_t; NULL = the handle's stream) without waiting.
 * Every call on one GPU shares that GPU's array: calls that overlap in time
 * on one GPU (other streams, other threads, other handles) race. */
int solar_compute_rhs_noalloc(solar_ctx* ctx, const void* q, void* q_rhs, int flags, void* stream);

/* FILTER_EXP(q, qf) (filter.f90:9-48) on host arrays; does not touch the state. */
int solar_filter_exp(solar_ctx* ctx, const double* q, double* qf);

/* Ghost strips each panel receives (the MPI panel exchange of DDX/DDY,
 * derivative.f90:55-133,374-452, done by the device gather):
 *   scalar f(n,n,nr,6) -> gx(3,n,nr,2,6) [DDX fp,fn], gy(n,3,nr,2,6) [DDY fp,fn];
 *   vector (v_xi, v_eta) -> receiver-basis v_xi ghosts in gx and v_eta
 *   ghosts in gy (COMPUTE_DIVERGENCE, grad.f90:95-122).
 * Host arrays; for tests of the exchange. */
int solar_scalar_ghosts(solar_ctx* ctx, const double* f, double* gx, double* gy);
int solar_vector_ghosts(solar_ctx* ctx, const double* vxi, const double* veta, double* gx, double* gy);

/* Non-finite scan of the device state: returns SOLAR_ERR_NONFINITE if any
 * element is NaN or Inf, SOLAR_OK otherwise.  (The reference has no check.) */
int solar_check_finite(solar_ctx* ctx);

/* RMS of v_r over all points, the diagnostic of cube.f90:77 (auxiliary.f90:100-125,
 * without its factor-of-6 slip at :122). */
int solar_norm_vr(solar_ctx* ctx, double* rms);

/* Average device time (ms) of each kernel class over the last
 * solar_rk4_async/solar_rk4 call, measured with HIP events on the handle's
 * stream: out[0] stage kernel, out[1] stage gather, out[2] filter kernel,
 * out[3] filter gather; out[4] total per step.  Needs SOLAR_PROFILE=1 in the
 * environment at solar_init; returns SOLAR_ERR_ARG otherwise. */
int solar_kernel_times(solar_ctx* ctx, double* out);

/* Bytes of device memory held by the handle. */
int solar_device_bytes(solar_ctx* ctx, size_t* bytes);

/* Algorithmic HBM bytes per grid point of each launch of this handle's RK4
 * step (state reads and writes, without halos, ghosts or tables):
 * out[0..4] the five stages, out[5] the filter.  With a 64-bit state the
 * step stores q after stages 2, 4 and 5 only (stages 2 and 4 rebuild their
 * input from q and the accumulator): 80, 160, 120, 160, 120 and 80 bytes,
 * 720 per step; with SOLAR_STEP_STORE_ALL=1 at solar_init, or a 32-bit
 * state, every stage stores q: 120, 160, 160, 160, 120, 80 (800; half that
 * for 32-bit).  No counterpart in the reference (for benchmarks). */
int solar_algorithmic_bytes(solar_ctx* ctx, double* out);

/* Sensitivities (discrete adjoint).  No counterpart in the reference, which
 * has no adjoint: these serve inversion and data-assimilation callers of the
 * same RK4 step.  All need a 64-bit state handle. ----------------------- */

/* Replaces the radial c^2(r) of solar_init by a full field c2(n,n,nr,6) in
 * COMPUTE_RHS's pressure equation: the reference's c2(i,j,k) is already a 3-D
 * array (compute_rhs_acoustic.f90:134) that SOLAR_DATA happens to fill
 * radially.  rho0 and g0 stay radial.  NULL restores the radial profile.
 * The host array is copied. */
int solar_set_c2_field(solar_ctx* ctx, const double* c2);

/* Storage of the stage divergences the misfit gradient keeps between its
 * forward and reverse passes: div_bits = 64 (default; the gradient is the
 * exact discrete gradient to rounding) or 32 (half the memory per stored
 * step, so longer segments and less recompute; the gradient then carries the
 * FP32 rounding of div(v), about 1e-7 relative). */
int solar_set_adjoint_options(solar_ctx* ctx, int div_bits);

/* Divergence cache of solar_misfit_gradient with a given checkpoint spacing:
 * enable != 0 (the default) lets the forward pass also keep the stage
 * divergences of as many trailing segments as free device memory holds, so
 * that only the earlier segments are recomputed from their checkpoints; 0
 * keeps the last segment only (every earlier segment recomputed).  The
 * gradient is the same bits either way. */
int solar_set_adjoint_cache(solar_ctx* ctx, int enable);

/* Steps the last solar_misfit_gradient recomputed from checkpoints. */
int solar_adjoint_recomputed_steps(solar_ctx* ctx, int* steps);

/* Receivers: nrec points of the p' field given as (panel, k, j, i) 0-based
 * quadruples in pkji[4*nrec]. */
int solar_set_receivers(solar_ctx* ctx, int nrec, const int* pkji);

/* nsteps RK4 steps from the current state, sampling p' at the receivers after
 * every step: seis[nsteps][nrec]. */
int solar_record(solar_ctx* ctx, int nsteps, double dt, double* seis);

/* Transposes of the step's linear parts on host arrays q(n,n,nr,5,6):
 *   solar_adj_rhs    out = L^T w,  L = SPONGE . BC_RHS . COMPUTE_RHS
 *   solar_adj_filter out = F^T o,  F = FILTER_EXP
 *   solar_adj_step   out = S^T lam, S = one RK4 step (timestep.f90:41-86)
 * Each is exact to rounding: <L x, w> = <x, L^T w>.  Do not touch the state. */
int solar_adj_rhs(solar_ctx* ctx, const double* w, double* out);
int solar_adj_filter(solar_ctx* ctx, const double* o, double* out);
int solar_adj_step(solar_ctx* ctx, double dt, const double* lam, double* out);

/* Tangent-linear model of the receiver traces with respect to the c^2 field:
 * dseis[n][r] = d p'_n(r) / d c2 . dc2 for nsteps RK4 steps from q0 (the same
 * traces solar_record samples), run alongside the primal steps with the
 * -dc2 rho0 div(v) source in every stage.  With the misfit gradient it gives
 * the dot-product test sum 2 (seis - data) dseis = <dJ/dc2, dc2>.
 * Overwrites the handle's state. */
int solar_tangent_linear(solar_ctx* ctx, const double* q0, int nsteps, double dt, const double* dc2, double* dseis);

/* Misfit J = sum over steps n and receivers r of (p'_n(r) - data[n][r])^2,
 * p'_n the state after step n+1 of nsteps RK4 steps from q0, and its
 * gradient grad(n,n,nr,6) = dJ/dc2 with respect to the c^2 field in use
 * (radial profile or solar_set_c2_field).  Checkpoint/recompute: the steps
 * are cut into segments of checkpoint_every steps (<= 0: the longest that
 * fits in free device memory), the last segment ending at nsteps.  The
 * forward pass keeps the state at the start of every other segment and the
 * stage divergences div(v) of the last one (of the last ones that fit, with
 * the divergence cache); the reverse pass recomputes each earlier segment
 * once from its checkpoint, storing its divergences, and runs the adjoint
 * steps over it.  data[nsteps][nrec]; seis (may be NULL) receives
 * the modelled traces; *segment_used (may be NULL) the segment length.
 * Overwrites the handle's state.  Device memory beyond the handle's:
 * segment * 40 B per point for the divergences plus one state per earlier
 * segment.  No sum depends on scheduling (no atomics): the same inputs give
 * the same bits, whatever checkpoint_every. */
int solar_misfit_gradient(solar_ctx* ctx, const double* q0, int nsteps, double dt, const double* data,
                          int checkpoint_every, double* J, double* grad, double* seis, int* segment_used);

/* Step-adjacent output in the reference's on-disk formats. ------------------
 * The reference runs one MPI rank per panel at px = py = pr = 1: rank = panel
 * 0..5, zoneid = rank + 1 (init_mpi.f90:34-37).  Paths are relative to dir
 * (NULL or "" = the working directory, as the reference); missing directories
 * are created.  <ddd> = digits cue/100, (cue%100)/10, cue%10 (output.f90:121-127).
 *
 * WRITE_RESTART (output.f90:86-190): Restart/restart.<ddd>.<zoneid>, per zone
 * int32 nx, ny, nr, 5; real*8 time; int32 step; then q(nx,ny,nr,5) real*8 in
 * Fortran order from byte 28.  solar_read_restart is its inverse (the
 * reference has none) and checks the sizes.  WRITE_VISUAL (output.f90:192-240):
 * Visual/<ddd>/vis_<rank>, one unformatted sequential record (int32 length
 * markers) of q(:,:,o_rad,1:5), o_rad 1-based.  WRITE_RADIAL_TRACE
 * (output.f90:312-343): appends ((q(1,1,k,l), k=1..nr), l=1..5) to
 * trace_<zoneid> in format (1x, 5(ES13.5)).  The handle versions write the
 * device state (FP32 storage is widened to real*8). */
int solar_write_restart(solar_ctx* ctx, const char* dir, int cue, double time, int step);
int solar_read_restart(solar_ctx* ctx, const char* dir, int cue, double* time, int* step);
int solar_write_visual(solar_ctx* ctx, const char* dir, int cue, int o_rad);
int solar_write_radial_trace(solar_ctx* ctx, const char* dir);

/* The same on host arrays q(n,n,nr,5,6) (no device needed). */
int solar_restart_write(const char* dir, int n, int nr, const double* q, int cue, double time, int step);
int solar_restart_read(const char* dir, int n, int nr, double* q, int cue, double* time, int* step);
int solar_visual_write(const char* dir, int n, int nr, const double* q, int cue, int o_rad);
int solar_radial_trace_write(const char* dir, int n, int nr, const double* q);
/* WRITE_GRID (output.f90:242-267): Visual/grid/grid_<rank>, (1x, 7i5). */
int solar_grid_write(const char* dir, int n, int nr);
/* o_rad of SOLAR_DATA (ic_acoustic_wave.f90:128-133): the 1-based radial
 * index closest to obsheight, ties to the larger index. */
int solar_output_radius(int nr, const double* gr, double obsheight, int* o_rad);

/* The norm diagnostic of cube.f90:77 for field 0..4 of the device state:
 * reference_formula != 0 gives auxiliary.f90:100-125 as written,
 * (sum over all panels of f^2 / (nx*ny*nr) * 6)**0.5; 0 gives the RMS over
 * the 6 panels (solar_norm_vr for field 3). */
int solar_norm(solar_ctx* ctx, int field, int reference_formula, double* out);

/* Batched ensembles. -------------------------------------------------------
 * An ensemble holds independent simulations with uneven grids (each member is
 * one run of the reference's cube.f90 time loop, cube.f90:71-90, with its own
 * n, nr, background, sponge and dt) and steps all of them together: one RK4
 * step is one hipGraph replay of a fixed number of launches, each covering
 * every member (a tile work list per tile shape).  FP64 state. */
typedef struct solar_ens solar_ens;

/* nmembers members; n[m], nr[m] as solar_init; gr, c2, rho0, g0 are the
 * members' radial profiles laid end to end (member m's start at the sum of
 * nr over members < m); sponge_len[m], a_sponge[m]; dt[m] > 0 the member's
 * time step.  States start at zero. */
int solar_ens_create(solar_ens** ens, int nmembers, const int* n, const int* nr, const double* gr, const double* c2,
                     const double* rho0, const double* g0, const int* sponge_len, const double* a_sponge,
                     const double* dt);
int solar_ens_finalize(solar_ens* ens);

/* Times the candidate tile shapes and radial chunk lengths of the stage and
 * filter kernels for every class of members with equal (n, nr), reps
 * launches each, and keeps the fastest; keeps those choices only if a whole
 * step with them beats one with the default shapes; then picks one block per
 * work item or a persistent grid from the time of a whole step.  Clears
 * every state: call it before solar_ens_set_state. */
int solar_ens_autotune(solar_ens* ens, int reps);

/* Launch parameters of member m: six[] = stage TI, TJ, planes per chunk,
 * filter TI, TJ, planes per chunk.  set: stage shapes 64x8, 32x8, 16x16 and
 * filter shapes 32x16, 16x16, 32x8 dividing n. */
int solar_ens_tiles(solar_ens* ens, int m, int* six);
int solar_ens_set_tiles(solar_ens* ens, int m, const int* six);

/* State of member m, q(n,n,nr,5,6) as for solar_set_state. */
int solar_ens_set_state(solar_ens* ens, int m, const double* q);
int solar_ens_get_state(solar_ens* ens, int m, double* q);

/* nsteps RK4 steps of every member (each with its own dt). */
int solar_ens_rk4(solar_ens* ens, int nsteps);
int solar_ens_rk4_async(solar_ens* ens, int nsteps);
int solar_ens_sync(solar_ens* ens);

/* Grid points over all members, launches per step, device bytes held. */
int solar_ens_info(solar_ens* ens, size_t* points, int* launches_per_step, size_t* device_bytes);

/* Phases of the last solar_misfit_gradient, seconds: out[0] workspace
 * allocation (0 when the handle's workspace from an earlier call fits),
 * out[1] forward pass (host clock), out[2] segment recomputes and out[3]
 * adjoint steps (HIP events on the handle's stream).  The workspace
 * (checkpoints, divergences, traces) stays with the handle until
 * solar_finalize. */
int solar_adjoint_times(solar_ctx* ctx, double* out);

/* Description of the last error on this thread ("" if none). */
const char* solar_last_error(void);

/* Host-only helpers (no device needed). -------------------------------- */

/* Time-invariant tables exactly as the device uses them, for the table
 * parity test: scal[4] = dxi, deta, dchi, pi; n1[8*n] = gxi, geta, tanxi,
 * taneta, metric_C, metric_D, global_C, global_D; r1[3*nr] = gchi, drdchi,
 * gr; n2[4*n*n] = metric_delta, sqrt_metric_delta, XY, CD; interp[27*n] =
 * interpolant(-4:4, n, 3); ghost[24] = X(-l), X(l), Y(-l), Y(l), C(-l),
 * C(l), D(-l), D(l) for l = 1..3; sigma[nr] = sponge coefficients;
 * drho0dr[nr] = DDR(rho0). */
int solar_host_tables(int n, int nr, const double* gr, const double* c2, const double* rho0, const double* g0,
                      int sponge_len, double a_sponge, double* scal, double* n1, double* r1, double* n2,
                      double* interp, double* ghost, double* sigma, double* drho0dr);

/* Seeded inputs (the caller-side SOLAR_DATA / GENERATE_IC). */
int solar_background(int kind, int nr, double* gr, double* c2, double* rho0, double* g0);
int solar_panel_points(int n, double* xyz);
int solar_generate_ic(int kind, int n, int nr, const double* gr, const double* c2, uint64_t seed,
                      const double* params, double* q, double* info);

#ifdef __cplusplus
}
#endif
#endif /* SOLAR_AMD_H */
