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
#define SOLAR_RHS_SPONGE 2 /* apply SPONGE (sponge.f90:215-249) */

/*
 * Set-up of the time-invariant tables and device buffers.  Replaces
 * SETUP_COORDINATES_METRICS_ALL (metric.f90:8), SETUP_LAGRANGE_INTERPOLANT
 * (interpolation.f90:9), the partner-rank tables of INITIALIZE_MPI
 * (init_mpi.f90:164-304) and the background arrays SOLAR_DATA fills
 * (ic_acoustic_wave.f90:257-271, radial-only there, so passed as profiles).
 *   n          angular points per panel edge (nx = ny), multiple of 16
 *   nr         radial points, >= 12
 *   gr, c2, rho0, g0   radial profiles, length nr (nondimensional)
 *   sponge_len, a_sponge   sponge.f90:215,222 (k_sponge_length, a_sponge)
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

/* Sensitivities (discrete adjoint).  No counterpart in the reference, which
 * has no adjoint: these serve inversion and data-assimilation callers of the
 * same RK4 step.  All need a 64-bit state handle. ----------------------- */

/* Replaces the radial c^2(r) of solar_init by a full field c2(n,n,nr,6) in
 * COMPUTE_RHS's pressure equation: the reference's c2(i,j,k) is already a 3-D
 * array (compute_rhs_acoustic.f90:134) that SOLAR_DATA happens to fill
 * radially.  rho0 and g0 stay radial.  NULL restores the radial profile.
 * The host array is copied. */
int solar_set_c2_field(solar_ctx* ctx, const double* c2);

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

/* Misfit J = sum over steps n and receivers r of (p'_n(r) - data[n][r])^2,
 * p'_n the state after step n+1 of nsteps RK4 steps from q0, and its
 * gradient grad(n,n,nr,6) = dJ/dc2 with respect to the c^2 field in use
 * (radial profile or solar_set_c2_field).  Checkpoint/recompute: the steps
 * are cut into segments of checkpoint_every steps (<= 0: the longest that
 * fits in free device memory), the last segment ending at nsteps.  The
 * forward pass keeps the state at the start of every other segment and the
 * stage divergences div(v) of the last one; the reverse pass recomputes each
 * earlier segment once from its checkpoint, storing its divergences, and runs
 * the adjoint steps over it.  data[nsteps][nrec]; seis (may be NULL) receives
 * the modelled traces; *segment_used (may be NULL) the segment length.
 * Overwrites the handle's state.  Device memory beyond the handle's:
 * segment * 40 B per point for the divergences plus one state per earlier
 * segment. */
int solar_misfit_gradient(solar_ctx* ctx, const double* q0, int nsteps, double dt, const double* data,
                          int checkpoint_every, double* J, double* grad, double* seis, int* segment_used);

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
