// Device kernels of the acoustic RK4 hot path for gfx950 (MI355X), FP64.
//
// State layout in HBM: q(n,n,nr,5,6), i fastest, one n*n*nr block per field
// per panel.  All arithmetic follows the reference's expression order and is
// compiled with -ffp-contract=off, so results match the CPU restatement of
// the reference bit for bit (tables are built on the host, see
// solar_tables.cpp).
//
//   k_gather  : the panel exchange (derivative.f90:55-133,374-452) as an
//               on-device gather: 9-tap Lagrange interpolation along the
//               sender's edge rows (interpolation.f90:140-378), for vectors
//               with the basis conversion at the source points
//               (transform.f90:8-59) and the receiver-basis transform
//               (transform.f90:114-276).
//   k_stage   : one RK stage fused in a single pass over HBM: DDX/DDY/DDR
//               (derivative.f90), COMPUTE_DIVERGENCE / COMPUTE_GRADIENT
//               (grad.f90:13-136), the RHS (compute_rhs_acoustic.f90:38-149),
//               BC_RHS, SPONGE and the 2N update (timestep.f90:64-69).
//   k_filter  : FILTER_EXP (filter.f90:9-781): xi then eta filter through an
//               LDS plane tile, then the 11-tap radial filter from a
//               register window.
#include <hip/hip_runtime.h>

#include "solar_kernels.hpp"

namespace solar {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Source point (i,j) of along-edge index a on the strip of edge e at layer l
// (1..3): S rows l, N rows n-1-l, W columns l, E columns n-1-l (0-based),
// i.e. rows/columns 2..4 and N-1..N-3 of grad.f90:28-31.
__device__ __forceinline__ int strip_index(int e, int l, int a, int n) {
  switch (e) {
    case kS: return l * n + a;
    case kN: return (n - 1 - l) * n + a;
    case kW: return a * n + l;
    default: return a * n + (n - 1 - l);
  }
}

// One thread per ghost value (receiver panel P, edge e, layer l, along index
// m, plane k).  VEC: also the receiver-basis velocity ghost.
//   q      : state, the scalar is field `sfield`, the vector fields 1, 2
//   gx*, gy*: ghosts, x: [P][side][k][m][l], y: [P][side][k][l][m]
template <bool VEC>
__global__ __launch_bounds__(256) void k_gather(DevTables T, const double* __restrict__ q, int sfield,
                                                double* __restrict__ gxs, double* __restrict__ gys,
                                                double* __restrict__ gxv, double* __restrict__ gyv,
                                                long long field_stride_out) {
  const int n = T.n, nr = T.nr;
  const long long total = 72LL * n * nr;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int fld = blockIdx.y;  // extra field offset (filter gather: 0..4)
  const int m = (int)(t % n);
  t /= n;
  const int l = (int)(t % 3) + 1;
  t /= 3;
  const int e = (int)(t % 4);
  t /= 4;
  const int P = (int)(t % 6);
  const int k = (int)(t / 6);

  const int s = T.route[P * 4 + e].sender;
  const int se = T.route[P * 4 + e].sender_edge;
  const int mp = T.route[P * 4 + e].reverse ? n - 1 - m : m;
  const size_t plane = (size_t)n * n, nf = plane * nr;
  const double* w = T.gw + ((size_t)(l - 1) * n + mp) * 9;
  const double* sp = q + ((size_t)(s * 5 + sfield + fld) * nr + k) * plane;

  double ps = 0.0, ths = 0.0, phs = 0.0;
  const double* vx = q + ((size_t)(s * 5 + 1) * nr + k) * plane;
  const double* ve = vx + nf;
  const int ctype = s < 4 ? 0 : s - 3;
  const double* cv = T.conv + (size_t)ctype * plane * 6;
#pragma unroll
  for (int tp = 0; tp < 9; ++tp) {
    const int a = clampi(mp - 4 + tp, 0, n - 1);
    const int idx = strip_index(se, l, a, n);
    const double wt = w[tp];
    ps = ps + sp[idx] * wt;
    if (VEC) {
      const double* c = cv + (size_t)idx * 6;
      const double x = vx[idx], y = ve[idx];
      const double th = c[0] * (c[1] * x + c[2] * y);
      const double ph = c[3] * (c[4] * x + c[5] * y);
      ths = ths + th * wt;
      phs = phs + ph * wt;
    }
  }
  size_t o;
  const size_t fo = (size_t)fld * field_stride_out;
  if (e < 2) {
    o = fo + (((size_t)(P * 2 + e) * nr + k) * n + m) * 3 + (l - 1);
    gxs[o] = ps;
  } else {
    o = fo + (((size_t)(P * 2 + (e - 2)) * nr + k) * 3 + (l - 1)) * n + m;
    gys[o] = ps;
  }
  if (VEC) {
    const double* ab = T.trans + ((((size_t)s * 4 + se) * 3 + (l - 1)) * n + mp) * 2;
    const double g = ab[0] * ths + ab[1] * phs;
    if (e < 2)
      gxv[o] = g;
    else
      gyv[o] = g;
  }
}

// Block index remap so that consecutive logical blocks (neighbouring tiles,
// whose halos overlap) run on the same XCD and share its L2.  Hardware
// dispatches block b to XCD b % 8.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  if (nb % 8 != 0) return b;
  return (b % 8) * (nb / 8) + b / 8;
}

__device__ __forceinline__ double pick7(const double (&w)[7], int c) {
  // Select w[c] for a block-uniform c without dynamic register indexing.
  double v = w[0];
#pragma unroll
  for (int q = 1; q < 7; ++q) v = (c == q) ? w[q] : v;
  return v;
}

// One RK stage (or one RHS evaluation when rhs_mode is set).
//   KIND 0: first stage (qw taken as 0, written), 1: middle stage (qw read and
//   written), 2: last stage (qw read, not written back).
template <int TI, int TJ, int KIND>
__global__ __launch_bounds__(TI* TJ) void k_stage(DevTables T, const double* __restrict__ qin,
                                                   double* __restrict__ qout, double* __restrict__ qw,
                                                   const double* __restrict__ gxp, const double* __restrict__ gyp,
                                                   const double* __restrict__ gxv, const double* __restrict__ gyv,
                                                   double alpha, double beta, double dt, int kchunk, int nkch,
                                                   int rhs_mode) {
  __shared__ double sp[TJ + 6][TI + 6];
  __shared__ double svx[TJ][TI + 6];
  __shared__ double sve[TJ + 6][TI];

  const int n = T.n, nr = T.nr;
  const int nti = n / TI, ntj = n / TJ;
  const int nb = nti * ntj * 6 * nkch;
  int b = xcd_remap(blockIdx.x, nb);
  const int bi = b % nti;
  b /= nti;
  const int bj = b % ntj;
  b /= ntj;
  const int P = b % 6;
  const int ch = b / 6;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = bi * TI, j0 = bj * TJ;
  const int i = i0 + tx, j = j0 + ty;
  const int k0 = ch * kchunk;
  const int k1 = min(k0 + kchunk, nr);
  if (k0 >= nr) return;

  const size_t plane = (size_t)n * n, nf = plane * nr;
  const size_t ij = (size_t)j * n + i;
  const double* f0 = qin + (size_t)(P * 5) * nf;  // rho'
  const double* f1 = f0 + nf;                     // v_xi
  const double* f2 = f1 + nf;                     // v_eta
  const double* f3 = f2 + nf;                     // v_r
  const double* f4 = f3 + nf;                     // p'
  double* o0 = qout + (size_t)(P * 5) * nf;
  double* w0 = qw + (size_t)(P * 5) * nf;

  const double Ci = T.ci[i], Dj = T.dj[j];
  const double dc = T.dc[ij], tdx = T.tdx[ij], tde = T.tde[ij], xy = T.xy[ij];

  // Ghost rows of this panel.
  const size_t gx_side = (size_t)nr * n * 3, gy_side = (size_t)nr * 3 * n;
  const double* gxpW = gxp + (size_t)(P * 2) * gx_side;
  const double* gxvW = gxv + (size_t)(P * 2) * gx_side;
  const double* gypS = gyp + (size_t)(P * 2) * gy_side;
  const double* gyvS = gyv + (size_t)(P * 2) * gy_side;

  // Radial windows of p' and v_r over planes kb..kb+6.
  double pw[7], rw[7];
  int kb = clampi(k0 - 3, 0, nr - 7);
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    pw[q] = f4[(size_t)(kb + q) * plane + ij];
    rw[q] = f3[(size_t)(kb + q) * plane + ij];
  }

  for (int k = k0; k < k1; ++k) {
    const int want = clampi(k - 3, 0, nr - 7);
    if (want != kb) {  // slides by exactly one plane
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        pw[q] = pw[q + 1];
        rw[q] = rw[q + 1];
      }
      kb = want;
      pw[6] = f4[(size_t)(kb + 6) * plane + ij];
      rw[6] = f3[(size_t)(kb + 6) * plane + ij];
    }
    const int c = k - kb;
    const double pc = pick7(pw, c);
    const double rc = pick7(rw, c);
    const size_t kp = (size_t)k * plane;
    const double vx = f1[kp + ij];
    const double ve = f2[kp + ij];

    // Plane tile with halos; panel-edge halos come from the gathered ghosts.
    sp[ty + 3][tx + 3] = pc;
    svx[ty][tx + 3] = vx;
    sve[ty + 3][tx] = ve;
    if (tx < 3) {
      const int ii = i0 - 3 + tx;
      if (ii >= 0) {
        sp[ty + 3][tx] = f4[kp + (size_t)j * n + ii];
        svx[ty][tx] = f1[kp + (size_t)j * n + ii];
      } else {
        const size_t g = ((size_t)k * n + j) * 3 + (-ii - 1);
        sp[ty + 3][tx] = gxpW[g];
        svx[ty][tx] = gxvW[g];
      }
    }
    if (tx >= TI - 3) {
      const int ii = i + 3;
      if (ii < n) {
        sp[ty + 3][tx + 6] = f4[kp + (size_t)j * n + ii];
        svx[ty][tx + 6] = f1[kp + (size_t)j * n + ii];
      } else {
        const size_t g = gx_side + ((size_t)k * n + j) * 3 + (ii - n);
        sp[ty + 3][tx + 6] = gxpW[g];
        svx[ty][tx + 6] = gxvW[g];
      }
    }
    if (ty < 3) {
      const int jj = j0 - 3 + ty;
      if (jj >= 0) {
        sp[ty][tx + 3] = f4[kp + (size_t)jj * n + i];
        sve[ty][tx] = f2[kp + (size_t)jj * n + i];
      } else {
        const size_t g = ((size_t)k * 3 + (-jj - 1)) * n + i;
        sp[ty][tx + 3] = gypS[g];
        sve[ty][tx] = gyvS[g];
      }
    }
    if (ty >= TJ - 3) {
      const int jj = j + 3;
      if (jj < n) {
        sp[ty + 6][tx + 3] = f4[kp + (size_t)jj * n + i];
        sve[ty + 6][tx] = f2[kp + (size_t)jj * n + i];
      } else {
        const size_t g = gy_side + ((size_t)k * 3 + (jj - n)) * n + i;
        sp[ty + 6][tx + 3] = gypS[g];
        sve[ty + 6][tx] = gyvS[g];
      }
    }
    __syncthreads();

    // 7-point DRP stencils, ascending taps from 0 (derivative.f90:137-148,
    // 456-467, 656-668, 704-754).
    double dvx = 0.0, dve = 0.0, dpx = 0.0, dpe = 0.0, dvr = 0.0, dpr = 0.0;
    const RadialRow& row = T.rad[k];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      dvx = dvx + T.fcx[q] * svx[ty][tx + q];
      dve = dve + T.fcy[q] * sve[ty + q][tx];
      dpx = dpx + T.fcx[q] * sp[ty + 3][tx + q];
      dpe = dpe + T.fcy[q] * sp[ty + q][tx + 3];
      dvr = dvr + row.w[q] * rw[q];
      dpr = dpr + row.w[q] * pw[q];
    }

    // grad.f90:125-133 and :49-57.
    const double div =
        (row.inv_r * ((dc * (dvx / Ci + dve / Dj) - tdx * vx) - tde * ve) + dvr) + (2.0 * rc) / row.r;
    const double gxi = row.inv_r * (Dj * dpx + (xy * dpe) / Dj);
    const double get = row.inv_r * (Ci * dpe + (xy * dpx) / Ci);

    // compute_rhs_acoustic.f90:38-149 (forcing omitted).
    const double q1 = f0[kp + ij];
    double r1 = -(row.rho0 * div);
    r1 = r1 - rc * row.drho0dr;
    double r2 = -(gxi / row.rho0);
    double r3 = -(get / row.rho0);
    double r4 = -(dpr / row.rho0);
    r4 = r4 - (q1 * row.g0) / row.rho0;
    double r5 = -(row.c2rho0 * div);
    r5 = r5 + row.rho0g0 * rc;

    const bool do_bc = rhs_mode ? (rhs_mode & 1) : true;
    const bool do_sp = rhs_mode ? (rhs_mode & 2) : true;
    if (do_bc && (k == 0 || k == nr - 1)) {  // bc_acoustic_wave.f90:14-34
      r1 = 0.0; r2 = 0.0; r3 = 0.0; r4 = 0.0; r5 = 0.0;
    }
    if (do_sp) {  // sponge.f90:243-245, after BC_RHS (timestep.f90:65-66)
      const double sg = row.sigma;
      r1 = r1 - sg * (q1 - 0.0);
      r4 = r4 - sg * (rc - 0.0);
      r5 = r5 - sg * (pc - 0.0);
    }

    const size_t x = kp + ij;
    if (rhs_mode) {
      o0[x] = r1;
      o0[nf + x] = r2;
      o0[2 * nf + x] = r3;
      o0[3 * nf + x] = r4;
      o0[4 * nf + x] = r5;
    } else {
      // timestep.f90:68-69: qw = alpha*qw + dt*rhs ; q = q + beta*qw
      const double rr[5] = {r1, r2, r3, r4, r5};
      const double qq[5] = {q1, vx, ve, rc, pc};
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        const double wold = (KIND == 0) ? 0.0 : w0[f * nf + x];
        const double wn = alpha * wold + dt * rr[f];
        o0[f * nf + x] = qq[f] + beta * wn;
        if (KIND != 2) w0[f * nf + x] = wn;
      }
    }
    __syncthreads();
  }
}

// FILTER_EXP for one field of one panel over a chunk of radial planes.
//   fgx/fgy: ghosts of the unfiltered field (filter.f90:25-28), layouts as in
//   k_gather, one block of 2*6 sides per field.
template <int TI, int TJ>
__global__ __launch_bounds__(TI* TJ) void k_filter(DevTables T, const double* __restrict__ qin,
                                                    double* __restrict__ qout, const double* __restrict__ fgx,
                                                    const double* __restrict__ fgy, int kchunk, int nkch) {
  __shared__ double sq[TJ + 6][TI + 6];
  __shared__ double sx[TJ + 6][TI];

  const int n = T.n, nr = T.nr;
  const int nti = n / TI, ntj = n / TJ;
  const int nb = nti * ntj * 30 * nkch;
  int b = xcd_remap(blockIdx.x, nb);
  const int bi = b % nti;
  b /= nti;
  const int bj = b % ntj;
  b /= ntj;
  const int PF = b % 30;  // panel*5 + field
  const int ch = b / 30;
  const int P = PF / 5, F = PF % 5;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = bi * TI, j0 = bj * TJ;
  const int i = i0 + tx, j = j0 + ty;
  const int k0 = ch * kchunk;
  const int k1 = min(k0 + kchunk, nr);
  if (k0 >= nr) return;

  const size_t plane = (size_t)n * n, nf = plane * nr;
  const double* g = qin + (size_t)PF * nf;
  double* o = qout + (size_t)PF * nf;
  const size_t gx_side = (size_t)nr * n * 3, gy_side = (size_t)nr * 3 * n;
  const size_t fstride = (size_t)12 * gx_side;  // one field's ghosts (6 panels x 2 sides)
  const double* gxW = fgx + F * fstride + (size_t)(P * 2) * gx_side;
  const double* gyS = fgy + F * fstride + (size_t)(P * 2) * gy_side;
  const size_t ij = (size_t)j * n + i;

  // filter.f90:652,726-777: the five planes at each end are zero.
  for (int k = k0; k < k1; ++k)
    if (k < kNstr || k >= nr - kNstr) o[(size_t)k * plane + ij] = 0.0;
  const int a = max(k0, kNstr), bnd = min(k1, nr - kNstr);
  if (a >= bnd) return;

  constexpr int NT = TI * TJ;
  const int tid = ty * TI + tx;
  double win[11];
#pragma unroll
  for (int q = 0; q < 11; ++q) win[q] = 0.0;

  for (int kk = a - kNstr; kk < bnd + kNstr; ++kk) {
    const size_t kp = (size_t)kk * plane;
    // Plane tile with halos (corners only where they lie inside the panel
    // in j, since out-of-panel rows are not xi-filtered).
    for (int e = tid; e < (TJ + 6) * (TI + 6); e += NT) {
      const int rr = e / (TI + 6), cc = e % (TI + 6);
      const int gj = j0 - 3 + rr, gi = i0 - 3 + cc;
      double v = 0.0;
      const bool jin = gj >= 0 && gj < n, iin = gi >= 0 && gi < n;
      if (jin && iin) {
        v = g[kp + (size_t)gj * n + gi];
      } else if (jin) {
        v = (gi < 0) ? gxW[((size_t)kk * n + gj) * 3 + (-gi - 1)]
                     : gxW[gx_side + ((size_t)kk * n + gj) * 3 + (gi - n)];
      } else if (iin) {
        v = (gj < 0) ? gyS[((size_t)kk * 3 + (-gj - 1)) * n + gi]
                     : gyS[gy_side + ((size_t)kk * 3 + (gj - n)) * n + gi];
      }
      sq[rr][cc] = v;
    }
    __syncthreads();
    // FILTER_XI on in-panel rows (filter.f90:176-315); rows beyond the panel
    // edge keep the unfiltered ghost, which FILTER_ETA uses (filter.f90:43).
    for (int rr = ty; rr < TJ + 6; rr += TJ) {
      const int gj = j0 - 3 + rr;
      double v;
      if (gj >= 0 && gj < n) {
        v = 0.0;
#pragma unroll
        for (int q = 0; q < 7; ++q) v = v + T.filt[q] * sq[rr][tx + q];
      } else {
        v = sq[rr][tx + 3];
      }
      sx[rr][tx] = v;
    }
    __syncthreads();
    // FILTER_ETA (filter.f90:448-629).
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < 7; ++q) v = v + T.filt[q] * sx[ty + q][tx];
#pragma unroll
    for (int q = 0; q < 10; ++q) win[q] = win[q + 1];
    win[10] = v;
    const int kout = kk - kNstr;
    if (kout >= a) {  // FILTER_R interior (filter.f90:682-693)
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < 11; ++q) s = s + T.filtr[q] * win[q];
      o[(size_t)kout * plane + ij] = s;
    }
    __syncthreads();
  }
}

// Non-finite scan and v_r sum of squares; acc[0] counts non-finite values,
// acc[1] accumulates squares.
__global__ void k_scan(const double* __restrict__ q, size_t total, size_t nf, double* acc, int mode) {
  double s = 0.0;
  for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (size_t)gridDim.x * blockDim.x) {
    if (mode == 0) {
      const double v = q[x];
      s += (isfinite(v) ? 0.0 : 1.0);
    } else {
      // v_r of panel p: field 3
      const size_t p = x / nf, r = x % nf;
      const double v = q[(p * 5 + 3) * nf + r];
      s += v * v;
    }
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(&acc[mode], s);
}

__global__ void k_zero(double* p, int count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) p[t] = 0.0;
}

// Gathers field slots of a staging state buffer into place: dst[(p*5+dfield)]
// <- src[p] for the ghost test entry points.
__global__ void k_place(const double* __restrict__ src, double* __restrict__ dst, size_t nf, int dfield) {
  const size_t total = nf * 6;
  for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (size_t)gridDim.x * blockDim.x) {
    const size_t p = x / nf, r = x % nf;
    dst[(p * 5 + dfield) * nf + r] = src[x];
  }
}

// ---------------------------------------------------------------- launchers

static int pick_ti(int n) { return (n % 64 == 0) ? 64 : ((n % 32 == 0) ? 32 : 16); }

int chunks_for(int tiles, int nr, int target_blocks, int min_chunk) {
  int nk = (target_blocks + tiles - 1) / tiles;
  const int maxk = nr / min_chunk > 0 ? nr / min_chunk : 1;
  if (nk > maxk) nk = maxk;
  if (nk < 1) nk = 1;
  return nk;
}

hipError_t launch_gather(const DevTables& T, const double* q, bool vec, int sfield, int nfields, double* gxs,
                         double* gys, double* gxv, double* gyv, hipStream_t st) {
  const long long total = 72LL * T.n * T.nr;
  const long long field_stride = 12LL * T.nr * T.n * 3;
  dim3 grid((unsigned)((total + 255) / 256), (unsigned)nfields);
  if (vec)
    hipLaunchKernelGGL(k_gather<true>, grid, dim3(256), 0, st, T, q, sfield, gxs, gys, gxv, gyv, field_stride);
  else
    hipLaunchKernelGGL(k_gather<false>, grid, dim3(256), 0, st, T, q, sfield, gxs, gys, gxv, gyv, field_stride);
  return hipGetLastError();
}

template <int TI, int TJ>
static hipError_t launch_stage_t(const DevTables& T, const double* qin, double* qout, double* qw, const double* gxp,
                                 const double* gyp, const double* gxv, const double* gyv, int stage, double dt,
                                 int rhs_mode, hipStream_t st, const LaunchShape& ls) {
  const int tiles = (T.n / TI) * (T.n / TJ) * 6;
  const int nkch = ls.stage_chunks > 0 ? ls.stage_chunks : chunks_for(tiles, T.nr, 2048, 32);
  const int kchunk = (T.nr + nkch - 1) / nkch;
  dim3 grid(tiles * nkch), block(TI, TJ);
  const double al = kAlphaRk4[stage < 0 ? 0 : stage], be = kBetaRk4[stage < 0 ? 0 : stage];
  if (rhs_mode || stage == 0)
    hipLaunchKernelGGL((k_stage<TI, TJ, 0>), grid, block, 0, st, T, qin, qout, qw, gxp, gyp, gxv, gyv, al, be, dt,
                       kchunk, nkch, rhs_mode);
  else if (stage < 4)
    hipLaunchKernelGGL((k_stage<TI, TJ, 1>), grid, block, 0, st, T, qin, qout, qw, gxp, gyp, gxv, gyv, al, be, dt,
                       kchunk, nkch, rhs_mode);
  else
    hipLaunchKernelGGL((k_stage<TI, TJ, 2>), grid, block, 0, st, T, qin, qout, qw, gxp, gyp, gxv, gyv, al, be, dt,
                       kchunk, nkch, rhs_mode);
  return hipGetLastError();
}

hipError_t launch_stage(const DevTables& T, const double* qin, double* qout, double* qw, const double* gxp,
                        const double* gyp, const double* gxv, const double* gyv, int stage, double dt, int rhs_mode,
                        hipStream_t st, const LaunchShape& ls) {
  switch (pick_ti(T.n)) {
    case 64: return launch_stage_t<64, 4>(T, qin, qout, qw, gxp, gyp, gxv, gyv, stage, dt, rhs_mode, st, ls);
    case 32: return launch_stage_t<32, 8>(T, qin, qout, qw, gxp, gyp, gxv, gyv, stage, dt, rhs_mode, st, ls);
    default: return launch_stage_t<16, 16>(T, qin, qout, qw, gxp, gyp, gxv, gyv, stage, dt, rhs_mode, st, ls);
  }
}

template <int TI, int TJ>
static hipError_t launch_filter_t(const DevTables& T, const double* qin, double* qout, const double* fgx,
                                  const double* fgy, hipStream_t st, const LaunchShape& ls) {
  const int tiles = (T.n / TI) * (T.n / TJ) * 30;
  const int nkch = ls.filter_chunks > 0 ? ls.filter_chunks : chunks_for(tiles, T.nr, 2048, 48);
  const int kchunk = (T.nr + nkch - 1) / nkch;
  dim3 grid(tiles * nkch), block(TI, TJ);
  hipLaunchKernelGGL((k_filter<TI, TJ>), grid, block, 0, st, T, qin, qout, fgx, fgy, kchunk, nkch);
  return hipGetLastError();
}

hipError_t launch_filter(const DevTables& T, const double* qin, double* qout, const double* fgx, const double* fgy,
                         hipStream_t st, const LaunchShape& ls) {
  switch (pick_ti(T.n)) {
    case 64: return launch_filter_t<64, 4>(T, qin, qout, fgx, fgy, st, ls);
    case 32: return launch_filter_t<32, 8>(T, qin, qout, fgx, fgy, st, ls);
    default: return launch_filter_t<16, 16>(T, qin, qout, fgx, fgy, st, ls);
  }
}

hipError_t launch_scan(const double* q, size_t nf, double* acc, int mode, hipStream_t st) {
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, st, acc, 2);
  const size_t total = mode == 0 ? nf * 30 : nf * 6;
  hipLaunchKernelGGL(k_scan, dim3(1024), dim3(256), 0, st, q, total, nf, acc, mode);
  return hipGetLastError();
}

hipError_t launch_place(const double* src, double* dst, size_t nf, int dfield, hipStream_t st) {
  hipLaunchKernelGGL(k_place, dim3(1024), dim3(256), 0, st, src, dst, nf, dfield);
  return hipGetLastError();
}

}  // namespace solar
