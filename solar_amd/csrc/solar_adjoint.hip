// Discrete adjoint of the acoustic RK4 step, FP64, gfx950.
//
// The forward step is linear in the state q: one stage computes
//   r = L(y) = SPONGE(BC_RHS(COMPUTE_RHS(y)))      (compute_rhs_acoustic.f90,
//   w = alpha_s w + dt r,  y = y + beta_s w         bc_acoustic_wave.f90,
//                                                  sponge.f90, timestep.f90:64-69)
// and the step ends with q = FILTER_EXP(y) (timestep.f90:73-75).  These
// kernels apply the exact transposes:
//   k_adj_stage    one reverse RK stage: Ybar += dt * L^T(W), the next W, and
//                  the sound-speed gradient  c2bar += dt * (-rho0 div) * W5.
//                  L^T = transposed DRP stencils (zero extension; the radial
//                  Tam-Dong rows transposed) applied to the transposed
//                  pointwise terms; edge tiles emit the adjoints of the ghost
//                  values they would have read.
//   k_adj_gathert  the transpose of the panel gather, as a gather over the
//                  sender's strips: ghost adjoints through the receiver-basis
//                  transform, the transposed 9-point Lagrange rows and the
//                  (xi,eta)->(theta,phi) conversion into the sender panel's
//                  near-edge rows and columns.  S/N strips, then W/E strips:
//                  strips of one pass share no point, so no add is atomic
//                  and the results do not depend on scheduling.
//   k_adj_filter   FILTER_R^T of the per-plane FILTER_XI^T FILTER_ETA^T.
//   k_adj_filter_ghosts  the per-plane ghost adjoints of both angular
//                  filter passes (the eta-pass ghosts come from the
//                  unfiltered field, filter.f90:25-28,43).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "solar_adjoint.hpp"
#include "solar_device.hpp"

// The library is built with -ffp-contract=off, so that the forward kernels
// round every product and sum as the reference does.  The adjoint has no
// reference bits to match: it is checked by transpose identities (to 1e-13
// of |Ax||y|), the tangent-linear dot product and finite differences.  Its
// kernels fuse products into sums (one rounding per multiply-add), which is
// as deterministic as the separate form: the gradient still repeats bit for
// bit across runs, checkpoint spacings and cache settings.
#pragma clang fp contract(fast)

namespace solar {

namespace {

__device__ __forceinline__ int strip_index_a(int e, int l, int a, int n) {
  switch (e) {
    case kS: return l * n + a;
    case kN: return (n - 1 - l) * n + a;
    case kW: return a * n + l;
    default: return a * n + (n - 1 - l);
  }
}

// Base plane of the DDR stencil of plane k (derivative.f90:656-754).
__device__ __forceinline__ int ddr_base(int k, int nr) { return k < 3 ? 0 : (k >= nr - 3 ? nr - 7 : k - 3); }

struct AdjU {
  double uxv, uyv, urv, uxp, uyp, urp;  // inputs of the transposed DDX/DDY/DDR
  double divb, b1, b4, b5;              // divergence adjoint and BC-masked W
};

// Transposed pointwise part of COMPUTE_RHS at one point (BC_RHS^T applied to
// W first: BC zeroes the end planes of the RHS).  The metric quotients come
// from tables (the transpose needs no bit-for-bit agreement with the forward
// expression order, only the same operator up to rounding): Ci = C(i),
// Dj = D(j), mq = (dc/C, dc/D, XY/C, XY/D) at the point.  Row is a radial row
// in any address space.
template <typename Row>
__device__ __forceinline__ void adj_point(const Row& row, bool bc, double Ci, double Dj, const double4& mq,
                                          double c2rho0, double w1, double w2, double w3, double w4, double w5,
                                          AdjU& u) {
  const double b1 = bc ? 0.0 : w1, b2 = bc ? 0.0 : w2, b3 = bc ? 0.0 : w3, b4 = bc ? 0.0 : w4,
               b5 = bc ? 0.0 : w5;
  // r1 = -rho0 div - v_r drho0dr ; r5 = -c2rho0 div + rho0 g0 v_r
  const double divb = -row.rho0 * b1 - c2rho0 * b5;
  // r2..r4 = -grad/rho0
  const double gxb = -b2 * row.inv_rho0, geb = -b3 * row.inv_rho0, grb = -b4 * row.inv_rho0;
  // div = inv_r*(dc*(Dx vxi/C + Dy veta/D) - tdx vxi - tde veta) + Dr vr + 2 vr/r
  u.uxv = row.inv_r * mq.x * divb;
  u.uyv = row.inv_r * mq.y * divb;
  u.urv = divb;
  // gxi = inv_r*(D Dx p + XY Dy p/D), geta = inv_r*(C Dy p + XY Dx p/C), gr = Dr p
  u.uxp = row.inv_r * (Dj * gxb + mq.z * geb);
  u.uyp = row.inv_r * (mq.w * gxb + Ci * geb);
  u.urp = grb;
  u.divb = divb;
  u.b1 = b1;
  u.b4 = b4;
  u.b5 = b5;
}

// The radial inputs (u_rv, u_rp) alone: they need no metric.
template <typename Row>
__device__ __forceinline__ void adj_radial(const Row& row, bool bc, double c2rho0, double w1, double w4, double w5,
                                           double& uv, double& up) {
  const double b1 = bc ? 0.0 : w1, b4 = bc ? 0.0 : w4, b5 = bc ? 0.0 : w5;
  uv = -row.rho0 * b1 - c2rho0 * b5;
  up = -b4 * row.inv_rho0;
}

}  // namespace

// One reverse RK stage of the adjoint (or L^T alone in RHS mode).
//   W    : W^(s) = the adjoint of the stage's 2N accumulator (5 fields)
//   Yin  : Ybar^(s) (5 fields); Yout = Ybar^(s-1) = Yin + dt L^T W (may alias Yin)
//   Wout : W^(s-1) = alpha*W + beta*Yout when mode & kAdjWriteW
//   divs : div(v) of the forward stage input, for the gradient (mode & kAdjGrad)
//   gc2  : c^2 gradient accumulator (n,n,nr,6)
//   gxa_*, gya_* : dt-scaled ghost adjoints written by edge tiles
// In RHS mode (mode & kAdjRhsOnly) Yout = L^T W and dt must be 1.
// D32: divs holds FP32 divergences (a template parameter, so the plane
// prefetch holds a single load path).
template <int TI, int TJ, bool D32>
__global__ __launch_bounds__(TI* TJ) void k_adj_stage(DevTables T, const double* __restrict__ W,
                                                       const double* Yin, double* Yout,  // may alias
                                                       double* __restrict__ Wout, const double* __restrict__ divs,
                                                       double* __restrict__ gc2, const double* __restrict__ c2f,
                                                       double* __restrict__ gxa_p, double* __restrict__ gya_p,
                                                       double* __restrict__ gxa_v, double* __restrict__ gya_v,
                                                       double dt, double alpha, double beta, double wscale,
                                                       int kchunk, int nkch, int mode) {
  // Transposed-stencil inputs with i halos (x) and j halos (y); two buffers
  // alternating by plane, so one barrier per plane suffices.
  __shared__ double sxv[2][TJ][TI + 6], sxp[2][TJ][TI + 6];
  __shared__ double syv[2][TJ + 6][TI], syp[2][TJ + 6][TI];
  // DRP taps for the edge tiles' ghost adjoints, which index them per
  // thread: from LDS, a dynamic index into the kernel argument would be a
  // vector load that the plane loop waits on together with the prefetch
  __shared__ double sfc[2][7];
  constexpr int NT = TI * TJ, NHX = 6 * TJ, NH = 6 * TJ + 6 * TI;
  static_assert(NH <= NT, "one halo cell per thread at most");

  const int n = T.n, nr = T.nr;
  const int nti = n / TI, ntj = n / TJ;
  int b = xcd_remap(blockIdx.x, nti * ntj * 6 * nkch);
  const int bi = b % nti;
  b /= nti;
  const int bj = b % ntj;
  b /= ntj;
  const int P = b % 6;
  const int ch = b / 6;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TI + tx;
  const int i0 = bi * TI, j0 = bj * TJ;
  const int i = i0 + tx, j = j0 + ty;
  const int k0 = ch * kchunk;
  const int k1 = min(k0 + kchunk, nr);
  if (k0 >= nr) return;

  const size_t plane = (size_t)n * n, nf = plane * nr;
  const size_t ij = (size_t)j * n + i;
  // W^(s), or wscale * Ybar when the stage starts from Ybar alone (W == nullptr)
  const double* w0 = (W ? W : Yin) + (size_t)(P * 5) * nf;
  const double ws = W ? 1.0 : wscale;
  const CRadialRows rows = const_rows(T.rad);
  auto c2rho0_of = [&](const auto& row, double c2) { return c2f ? c2 * row.rho0 : row.c2rho0; };
  auto c2_at = [&](int k, size_t cell) { return c2f ? c2f[(size_t)P * nf + (size_t)k * plane + cell] : 0.0; };

  // Radial inputs (u_rv, u_rp) of plane k at this column, loaded on demand
  // (the window fill and the closure planes).
  auto radial_u = [&](int k, double& uv, double& up) {
    if (k < 0 || k >= nr) {
      uv = 0.0;
      up = 0.0;
      return;
    }
    const size_t c = (size_t)k * plane + ij;
    adj_radial(rows[k], k == 0 || k == nr - 1, c2rho0_of(rows[k], c2_at(k, ij)), ws * w0[c], ws * w0[3 * nf + c],
               ws * w0[4 * nf + c], uv, up);
  };
  // Radial window over planes k-3..k+3.
  double rv[7], rp[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) radial_u(k0 - 3 + q, rv[q], rp[q]);

  // The halo cell this thread fills (at most one per thread when NH <= NT):
  // x halos first (TJ rows x 6 columns), then y halos (6 rows x TI columns).
  // hcell < 0: beyond the panel edge, where the transposed stencil reads 0.
  const bool has_h = tid < NH;
  bool hx = false;
  int hrow = 0, hcol = 0, hi_ = 0, hj_ = 0;
  if (has_h) {
    if (tid < NHX) {
      hx = true;
      hrow = tid / 6;
      const int c6 = tid % 6;
      hcol = c6 < 3 ? c6 : TI + c6;
      hi_ = i0 - 3 + hcol;
      hj_ = j0 + hrow;
    } else {
      const int h = tid - NHX;
      const int r6 = h / TI;
      hcol = h % TI;
      hrow = r6 < 3 ? r6 : TJ + r6;
      hi_ = i0 + hcol;
      hj_ = j0 - 3 + hrow;
    }
  }
  const bool h_in = has_h && hi_ >= 0 && hi_ < n && hj_ >= 0 && hj_ < n;
  const size_t hcell = h_in ? (size_t)hj_ * n + hi_ : 0;

  // Loop-invariant metric of the own point and, for the halo cell, the three
  // factors its one transposed-stencil input needs: x halos take
  // (dc/C, D, XY/C) for u_xv, u_xp; y halos (dc/D, XY/D, C) for u_yv, u_yp.
  const double Ci = T.ci[i], Dj = T.dj[j];
  const double4 mq = *reinterpret_cast<const double4*>(T.adj_metric + ij * 4);
  const double tdx = T.tdx[ij], tde = T.tde[ij];
  double hA = 0.0, hB = 0.0, hC = 0.0;
  if (h_in) {
    const double4 hm = *reinterpret_cast<const double4*>(T.adj_metric + hcell * 4);
    hA = hx ? hm.x : hm.y;
    hB = hx ? T.dj[hj_] : hm.w;
    hC = hx ? hm.z : T.ci[hi_];
  }

  const double sig_dt = (mode & kAdjRhsOnly) ? 1.0 : dt;
  const size_t gx_side = (size_t)nr * n * 3, gy_side = (size_t)nr * 3 * n;

  // W values of one plane, loaded a plane ahead: own point, halo cell, and
  // the plane entering the radial window (k+3 for plane k); Ybar, div(v), the
  // c^2 gradient and c^2 come with them.  Every load is issued whatever the
  // mode (unused inputs read a fixed in-bounds address), so no branch sits
  // between the loads and the waits before their use count exactly what is
  // still in flight.
  const bool upd = !(mode & kAdjRhsOnly), grad = (mode & kAdjGrad) != 0;
  const double* yi = upd ? Yin + (size_t)(P * 5) * nf : w0;
  const double* dvb = grad ? divs : w0;
  const double* gcb = grad ? gc2 : w0;
  const double* c2b = c2f ? c2f + (size_t)P * nf : w0;
  struct AdjPlane {
    double w[5], hw[4], rw[3], y[5], dv, g, c2o, c2h, c2r;
    float dvf;  // D32: the divergence as loaded, widened where it is used (a conversion at
                // the load would wait on every load in flight)
  };
  auto load_plane = [&](int k, AdjPlane& L) {
    const size_t x = (size_t)k * plane + ij;
#pragma unroll
    for (int f = 0; f < 5; ++f) L.w[f] = w0[(size_t)f * nf + x];
#pragma unroll
    for (int f = 0; f < 5; ++f) L.y[f] = __builtin_nontemporal_load(yi + (size_t)f * nf + (upd ? x : 0));
    const size_t g = grad ? (size_t)P * nf + x : 0;
    if constexpr (D32)
      L.dvf = __builtin_nontemporal_load(reinterpret_cast<const float*>(dvb) + g);
    else
      L.dv = __builtin_nontemporal_load(dvb + g);
    L.g = gcb[g];
    const size_t hk = (size_t)k * plane + hcell;
    L.hw[0] = w0[hk];
    L.hw[1] = w0[nf + hk];
    L.hw[2] = w0[2 * nf + hk];
    L.hw[3] = w0[4 * nf + hk];
    const size_t ek = (size_t)min(k + 3, nr - 1) * plane + ij;
    L.rw[0] = w0[ek];
    L.rw[1] = w0[3 * nf + ek];
    L.rw[2] = w0[4 * nf + ek];
    L.c2o = c2b[c2f ? x : 0];
    L.c2h = c2b[c2f ? hk : 0];
    L.c2r = c2b[c2f ? ek : 0];
  };
  if (threadIdx.y == 0 && threadIdx.x < 7) {  // read after the first plane's barrier
    sfc[0][threadIdx.x] = T.fcx[threadIdx.x];
    sfc[1][threadIdx.x] = T.fcy[threadIdx.x];
  }
  AdjPlane nxt;
  load_plane(k0, nxt);

  for (int k = k0; k < k1; ++k) {
    const AdjPlane cur = nxt;
    load_plane(min(k + 1, k1 - 1), nxt);  // the last plane's prefetch is loaded and not used
    const auto& row = rows[k];
    const bool bck = k == 0 || k == nr - 1;
    if (k > k0) {
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        rv[q] = rv[q + 1];
        rp[q] = rp[q + 1];
      }
      const int ke = k + 3;
      if (ke < nr) {
        const auto& re = rows[ke];
        adj_radial(re, ke == nr - 1, c2rho0_of(re, cur.c2r), ws * cur.rw[0], ws * cur.rw[1], ws * cur.rw[2], rv[6],
                   rp[6]);
      } else {
        rv[6] = 0.0;
        rp[6] = 0.0;
      }
    }
    const size_t kp = (size_t)k * plane;
    const int bf = (k - k0) & 1;
    // own point
    const double w1 = ws * cur.w[0], w2 = ws * cur.w[1], w3 = ws * cur.w[2], w4 = ws * cur.w[3], w5 = ws * cur.w[4];
    AdjU u;
    adj_point(row, bck, Ci, Dj, mq, c2rho0_of(row, cur.c2o), w1, w2, w3, w4, w5, u);
    sxv[bf][ty][tx + 3] = u.uxv;
    sxp[bf][ty][tx + 3] = u.uxp;
    syv[bf][ty + 3][tx] = u.uyv;
    syp[bf][ty + 3][tx] = u.uyp;
    if (has_h) {
      double hv = 0.0, hp = 0.0;
      if (h_in) {  // the halo cell's one input: u_xv, u_xp (x halo) or u_yv, u_yp (y halo)
        const double b1 = bck ? 0.0 : ws * cur.hw[0], b2 = bck ? 0.0 : ws * cur.hw[1],
                     b3 = bck ? 0.0 : ws * cur.hw[2], b5 = bck ? 0.0 : ws * cur.hw[3];
        const double divb = -row.rho0 * b1 - c2rho0_of(row, cur.c2h) * b5;
        const double gxb = -b2 * row.inv_rho0, geb = -b3 * row.inv_rho0;
        hv = row.inv_r * hA * divb;
        hp = row.inv_r * (hB * gxb + hC * geb);
      }
      if (hx) {
        sxv[bf][hrow][hcol] = hv;
        sxp[bf][hrow][hcol] = hp;
      } else {
        syv[bf][hrow][hcol] = hv;
        syp[bf][hrow][hcol] = hp;
      }
    }
    __syncthreads();

    // Transposed stencils: (D^T u)(i) = sum_q c[6-q] u(i-3+q); one set of
    // coefficients for xi and eta, whose zero centre tap adds nothing (as in
    // the forward stage; both checked when the tables are uploaded).
    double dxv = 0.0, dxp = 0.0, dyv = 0.0, dyp = 0.0;
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      if (q == 3) continue;
      dxv += T.fcx[6 - q] * sxv[bf][ty][tx + q];
      dxp += T.fcx[6 - q] * sxp[bf][ty][tx + q];
      dyv += T.fcx[6 - q] * syv[bf][ty + q][tx];
      dyp += T.fcx[6 - q] * syp[bf][ty + q][tx];
    }
    // Transposed radial stencil: every row kk whose DDR stencil reads plane k.
    double drv = 0.0, drp = 0.0;
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      const int kk = k - 3 + q;
      if (kk >= 0 && kk < nr) {
        const int t = k - ddr_base(kk, nr);
        if (t >= 0 && t <= 6) {
          const double wt = rows[kk].w[t];
          drv += wt * rv[q];
          drp += wt * rp[q];
        }
      }
    }
    // the closure rows 0..2 and nr-3..nr-1 reach planes 0..6 and nr-7..nr-1
    // (derivative.f90:704-716,742-754): those beyond the window, on demand
    if (k >= 4 && k <= 6) {
      for (int kk = 0; kk < k - 3; ++kk) {
        double uv, up;
        radial_u(kk, uv, up);
        const double wt = rows[kk].w[k];
        drv += wt * uv;
        drp += wt * up;
      }
    }
    if (k >= nr - 7 && k <= nr - 5) {
      for (int kk = k + 4; kk < nr; ++kk) {
        double uv, up;
        radial_u(kk, uv, up);
        const double wt = rows[kk].w[k - (nr - 7)];
        drv += wt * uv;
        drp += wt * up;
      }
    }

    const double sg = row.sigma;
    double lt[5];
    lt[0] = -(row.g0 * row.inv_rho0) * u.b4 - sg * w1;
    lt[1] = -row.inv_r * tdx * u.divb + dxv;
    lt[2] = -row.inv_r * tde * u.divb + dyv;
    lt[3] = -row.drho0dr * u.b1 + row.rho0g0 * u.b5 + (2.0 * row.inv_r) * u.divb - sg * w4 + drv;
    lt[4] = -sg * w5 + dxp + dyp + drp;

    const size_t x = kp + ij;
    double* yo = Yout + (size_t)(P * 5) * nf;
    if (mode & kAdjRhsOnly) {
#pragma unroll
      for (int f = 0; f < 5; ++f) yo[(size_t)f * nf + x] = lt[f];
    } else {
      double* wo = Wout + (size_t)(P * 5) * nf;
      const double wv[5] = {w1, w2, w3, w4, w5};
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        const double yn = cur.y[f] + dt * lt[f];
        __builtin_nontemporal_store(yn, yo + (size_t)f * nf + x);
        if (mode & kAdjWriteW) __builtin_nontemporal_store(alpha * wv[f] + beta * yn, wo + (size_t)f * nf + x);
      }
      if (grad) {
        // r5 = -c^2 rho0 div + ...  ->  d r5 / d c^2 = -rho0 div (zero on the BC planes)
        const double dv = D32 ? static_cast<double>(cur.dvf) : cur.dv;
        gc2[(size_t)P * nf + x] = cur.g + dt * (-row.rho0 * dv) * u.b5;
      }
    }

    // Ghost adjoints of edge tiles: the receiver's DDX/DDY read ghost layer
    // l at i = -l (W), n-1+l (E), j = -l (S), n-1+l (N) with tap 3-l-i, etc.
    if (bi == 0 && tx < 3) {
      const int l = tx + 1;
      double gp = 0.0, gv = 0.0;
      for (int ii = 0; ii <= 3 - l; ++ii) {
        gp += sfc[0][3 - l - ii] * sxp[bf][ty][3 + ii];
        gv += sfc[0][3 - l - ii] * sxv[bf][ty][3 + ii];
      }
      const size_t o = (size_t)(P * 2) * gx_side + ((size_t)k * n + j) * 3 + (l - 1);
      gxa_p[o] = sig_dt * gp;
      gxa_v[o] = sig_dt * gv;
    }
    if (bi == nti - 1 && tx >= TI - 3) {
      const int l = tx - (TI - 3) + 1;
      double gp = 0.0, gv = 0.0;
      for (int ii = n - 4 + l; ii <= n - 1; ++ii) {
        gp += sfc[0][n + 2 + l - ii] * sxp[bf][ty][3 + ii - i0];
        gv += sfc[0][n + 2 + l - ii] * sxv[bf][ty][3 + ii - i0];
      }
      const size_t o = (size_t)(P * 2 + 1) * gx_side + ((size_t)k * n + j) * 3 + (l - 1);
      gxa_p[o] = sig_dt * gp;
      gxa_v[o] = sig_dt * gv;
    }
    if (bj == 0 && ty < 3) {
      const int l = ty + 1;
      double gp = 0.0, gv = 0.0;
      for (int jj = 0; jj <= 3 - l; ++jj) {
        gp += sfc[1][3 - l - jj] * syp[bf][3 + jj][tx];
        gv += sfc[1][3 - l - jj] * syv[bf][3 + jj][tx];
      }
      const size_t o = (size_t)(P * 2) * gy_side + ((size_t)k * 3 + (l - 1)) * n + i;
      gya_p[o] = sig_dt * gp;
      gya_v[o] = sig_dt * gv;
    }
    if (bj == ntj - 1 && ty >= TJ - 3) {
      const int l = ty - (TJ - 3) + 1;
      double gp = 0.0, gv = 0.0;
      for (int jj = n - 4 + l; jj <= n - 1; ++jj) {
        gp += sfc[1][n + 2 + l - jj] * syp[bf][3 + jj - j0][tx];
        gv += sfc[1][n + 2 + l - jj] * syv[bf][3 + jj - j0][tx];
      }
      const size_t o = (size_t)(P * 2 + 1) * gy_side + ((size_t)k * 3 + (l - 1)) * n + i;
      gya_p[o] = sig_dt * gp;
      gya_v[o] = sig_dt * gv;
    }
  }
}

// Transpose of k_gather, as a gather over the sender's strips: one block per
// (sender panel s, sender edge se, chunk of planes).  Per plane the
// receiver's ghost values of that strip (receiver from the inverse route)
// are first staged in LDS in the sender's along-edge order -- for vectors
// already through the receiver-basis transform, as theta and phi parts --
// then each strip point (layer l, along index a) sums the receiver stencils
// that read it (the transposed rows gt_*), applies the sender's
// (xi,eta)->(theta,phi) conversion transposed for vectors, and adds the
// result to Y (and beta times it to Wn).  pass 0 covers the S/N strips, pass
// 1 the W/E strips: strips of one pass never share a point, so the adds need
// no atomics; corner points get their two strips in turn.  The transposed
// rows, transforms and conversions do not depend on the plane: each thread
// holds its own in registers.  RADF applies FILTER_R^T (filter.f90:682-693)
// to the per-plane ghost adjoints from an 11-plane register window, one new
// plane per step.
// Threads per block: one strip point per thread up to 1024 (n <= 341);
// beyond, vector strips (whose per-point tables take ~70 registers) use
// three points per thread at 512 threads, scalar strips two at 768.
// The next plane's ghost values load while this plane is summed, in two
// alternating buffers (the loop is unrolled by two), so no register copy
// between planes waits on the loads in flight.  Each plane's read-modify-
// write inputs are issued first in its step: the stores wait for them
// without draining the ghost prefetch issued after them.  A thread
// past the strip (val false) takes strip point 0's indices and tables: its
// loads, LDS writes and stores repeat that point's, with the same values, so
// none of them needs a branch.  HASW: Wn is not null.
constexpr int gathert_threads_max(bool vec, int ni) { return ni == 1 ? 1024 : (vec ? 512 : gather_threads_max(ni)); }

template <bool VEC, bool RADF, int NI, bool HASW>
__global__ __launch_bounds__(gathert_threads_max(VEC, NI)) void k_adj_gathert(
    DevTables T, const double* __restrict__ gxs, const double* __restrict__ gys, const double* __restrict__ gxv,
    const double* __restrict__ gyv, double* __restrict__ Y, double* __restrict__ Wn, double beta, int sfield,
    long long field_stride_in, int pass) {
  // 2 buffers x [3 layers][n]: scalar; then theta, phi parts for vectors
  extern __shared__ double sg[];
  constexpr int NQ = VEC ? 3 : 1;
  const int n = T.n, nr = T.nr, strip = 3 * n, bstride = NQ * strip;
  const int kc = gather_kc(nr);
  const int b = blockIdx.x;
  const int k0 = (b / 12) * kc, k1 = min(k0 + kc, nr);
  const int s = (b % 12) / 2;
  const int se = (pass == 0 ? kS : kW) + (b & 1);
  const int fld = blockIdx.y;
  const int rv = T.inv_route[s * 4 + se];
  const int P = (rv & 0xff) >> 2, e = rv & 3, rev = rv >> 8;
  const size_t plane = (size_t)n * n, nf = plane * nr;
  const size_t fo = (size_t)fld * field_stride_in;
  const size_t kst = (size_t)3 * n;  // plane stride of both ghost layouts
  const auto gsrc = to_global((e < 2) ? gxs : gys);
  const auto vsrc = to_global((e < 2) ? (VEC ? gxv : gxs) : (VEC ? gyv : gys));
  const size_t gside = fo + (size_t)(P * 2 + (e < 2 ? e : e - 2)) * nr * kst;
  const bool col = (se == kW || se == kE);  // column strips: layer index fastest in memory
  const int ctype = s < 4 ? 0 : s - 3;
  const double* cv = T.conv + (size_t)ctype * plane * 6;

  // this thread's receiver ghosts (staging) and strip points (sums)
  size_t gofs[NI];
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int t = threadIdx.x + it * blockDim.x;
    gofs[it] = gside + (t < strip ? t : 0);
  }
  // Ghost inputs of one plane k (RADF: of plane k+5, the one entering the
  // window; zeroed where it is used when k+5 lies beyond the filtered planes).
  struct In {
    double g[NI], gv[NI];
  };
  auto load_in = [&](int k, In& d) {
    const int kk = min(k, k1 - 1);  // past the chunk: loaded and not used
    const int kg = RADF ? min(kk + 5, nr - 1 - kNstr) : kk;
#pragma unroll
    for (int it = 0; it < NI; ++it) {
      d.g[it] = gsrc[gofs[it] + (size_t)kg * kst];
      if (VEC) d.gv[it] = vsrc[gofs[it] + (size_t)kk * kst];
    }
  };
  // The first plane's ghosts load before the tables, so that the loop is
  // entered with as many operations issued after them as on its back edge
  // (the plane's stores): the wait for them at the top of the loop then
  // leaves the stores in flight.
  In cur, nxt;
  load_in(k0, nxt);
  int slot[NI], gmp[NI][kGtWidth], idx[NI];
  double ab[NI][2], gw[NI][kGtWidth], c[NI][6];
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int t = threadIdx.x + it * blockDim.x;
    const int tt = t < strip ? t : 0;
    // x ghosts [m][l] (l fastest), y ghosts [l][m]
    const int l = (e < 2) ? tt % 3 + 1 : tt / n + 1;
    const int m = (e < 2) ? tt / 3 : tt % n;
    const int mp = rev ? n - 1 - m : m;
    slot[it] = (l - 1) * n + mp;
    if (VEC) {
      const double* abp = T.trans + ((((size_t)s * 4 + se) * 3 + (l - 1)) * n + mp) * 2;
      ab[it][0] = abp[0];
      ab[it][1] = abp[1];
    }
    const int ls = col ? tt % 3 + 1 : tt / n + 1;
    const int a = col ? tt / 3 : tt % n;
    const size_t r0 = ((size_t)(ls - 1) * n + a) * kGtWidth;
#pragma unroll
    for (int q = 0; q < kGtWidth; ++q) {
      gmp[it][q] = (ls - 1) * n + T.gt_mp[r0 + q];
      gw[it][q] = T.gt_w[r0 + q];
    }
    idx[it] = strip_index_a(se, ls, a, n);
    if (VEC) {
#pragma unroll
      for (int x = 0; x < 6; ++x) c[it][x] = cv[(size_t)idx[it] * 6 + x];
    }
  }
  const auto ys = to_global(Y + (size_t)(s * 5 + sfield + fld) * nf);
  const auto y1 = to_global(Y + (size_t)(s * 5 + 1) * nf);
  const auto y2 = y1 + nf;
  // without Wn these alias Y and are never loaded or stored
  const auto ws = to_global(HASW ? Wn + (size_t)(s * 5 + sfield + fld) * nf : Y);
  const auto w1 = to_global(HASW ? Wn + (size_t)(s * 5 + 1) * nf : Y);
  const auto w2 = w1 + nf;

  // RADF: window of per-plane ghost adjoints over planes k-5..k+5 (zero
  // outside the filtered planes [5, nr-6], which FILTER_R leaves alone)
  double win[NI][11];
  if (RADF) {
#pragma unroll
    for (int it = 0; it < NI; ++it)
#pragma unroll
      for (int q = 0; q < 10; ++q) {
        const int kk = k0 - 5 + q;
        win[it][q + 1] = (kk >= kNstr && kk <= nr - 1 - kNstr) ? gsrc[gofs[it] + (size_t)kk * kst] : 0.0;
      }
  }
  auto step = [&](int k, In& cur, In& nxt) {
    const size_t kp = (size_t)k * plane;
    double yv[NI], wv[NI], y1v[NI], y2v[NI], w1v[NI], w2v[NI];
#pragma unroll
    for (int it = 0; it < NI; ++it) {
      const size_t x = kp + idx[it];
      yv[it] = ys[x];
      if (HASW) wv[it] = ws[x];
      if (VEC) {
        y1v[it] = y1[x];
        y2v[it] = y2[x];
        if (HASW) {
          w1v[it] = w1[x];
          w2v[it] = w2[x];
        }
      }
    }
    double* buf = sg + ((k - k0) & 1) * bstride;
#pragma unroll
    for (int it = 0; it < NI; ++it) {
      double gsv;
      if (RADF) {
#pragma unroll
        for (int q = 0; q < 10; ++q) win[it][q] = win[it][q + 1];
        win[it][10] = (k + 5 <= nr - 1 - kNstr) ? cur.g[it] : 0.0;
        gsv = 0.0;  // sum over kk = k-5..k+5 ascending: filtr[k - kk + 5] = filtr[10 - q]
#pragma unroll
        for (int q = 0; q < 11; ++q) {
          const int kk = k - 5 + q;
          if (kk >= kNstr && kk <= nr - 1 - kNstr) gsv += T.filtr[10 - q] * win[it][q];
        }
      } else {
        gsv = cur.g[it];
      }
      buf[slot[it]] = gsv;
      if (VEC) {
        buf[strip + slot[it]] = ab[it][0] * cur.gv[it];
        buf[2 * strip + slot[it]] = ab[it][1] * cur.gv[it];
      }
    }
    load_in(k + 1, nxt);
    __syncthreads();
#pragma unroll
    for (int it = 0; it < NI; ++it) {
      double as = 0.0, ath = 0.0, aph = 0.0;
#pragma unroll
      for (int q = 0; q < kGtWidth; ++q) {
        const double w = gw[it][q];
        as += w * buf[gmp[it][q]];
        if (VEC) {
          ath += w * buf[strip + gmp[it][q]];
          aph += w * buf[2 * strip + gmp[it][q]];
        }
      }
      const size_t x = kp + idx[it];
      ys[x] = yv[it] + as;
      if (HASW) ws[x] = wv[it] + beta * as;
      if (VEC) {
        const double ax = c[it][0] * c[it][1] * ath + c[it][3] * c[it][4] * aph;
        const double ay = c[it][0] * c[it][2] * ath + c[it][3] * c[it][5] * aph;
        y1[x] = y1v[it] + ax;
        y2[x] = y2v[it] + ay;
        if (HASW) {
          w1[x] = w1v[it] + beta * ax;
          w2[x] = w2v[it] + beta * ay;
        }
      }
    }
  };
  for (int k = k0; k < k1; ++k) {
    cur = nxt;
    step(k, cur, nxt);
  }
}

// FILTER_EXP^T = (FILTER_R FILTER_ETA FILTER_XI)^T.  FILTER_R acts on the
// radial index alone and the two angular filters (with their ghost exchange)
// on each plane alone, so the transpose is taken as FILTER_R^T applied to
// FILTER_XI^T FILTER_ETA^T of each plane (equal in exact arithmetic).
// Per plane kk in [5, nr-6] (FILTER_R reads no other plane's output,
// filter.f90:682-693): the (TJ+6)x(TI+6) tile of bbar (zero beyond the panel:
// the transposed stencils have no ghosts) goes to LDS, FILTER_ETA^T runs on
// every column, FILTER_XI^T on the in-panel rows; the result enters an
// 11-plane register window from which FILTER_R^T forms the output plane
// kk - 5.  The ghost adjoints of both angular passes come from
// k_adj_filter_ghosts.
// Per-plane ghost adjoints of FILTER_XI^T and FILTER_ETA^T (the receiver's
// filters read ghost layer l at i = -l (W), n-1+l (E), j = -l (S), n-1+l (N)
// with tap 3-l-i, etc.), for planes kk in [5, nr-6]; zero on the others
// (FILTER_R reads no other plane's output).  The eta ghosts come from bbar
// itself (FILTER_ETA reads unfiltered ghosts, filter.f90:43), the xi ghosts
// from abar = FILTER_ETA^T bbar on the edge columns.  One block per (panel,
// field, edge, chunk of planes): per plane the three edge rows (S/N) or
// columns (W/E) of bbar go to LDS, one value per thread, with the next
// plane's value loading meanwhile; each thread then forms one (layer, along
// index) output.
template <int NI>
__global__ __launch_bounds__(gather_threads_max(NI)) void k_adj_filter_ghosts(DevTables T,
                                                                               const double* __restrict__ bin,
                                                                               double* __restrict__ fgxa,
                                                                               double* __restrict__ fgya) {
  extern __shared__ double sl[];  // 2 buffers x 3 lines x n
  // filter taps, indexed per thread: from LDS (a dynamic index into the
  // kernel argument is a vector load, waited on with the prefetched plane)
  __shared__ double sft[7];
  const int n = T.n, nr = T.nr, strip = 3 * n;
  const int kc = gather_kc(nr);
  const int b = blockIdx.x;
  const int k0 = (b / 120) * kc, k1 = min(k0 + kc, nr);
  const int e = (b % 120) % 4, PF = (b % 120) / 4;
  const int P = PF / 5, F = PF % 5;
  const size_t plane = (size_t)n * n, nf = plane * nr;
  const double* g = bin + (size_t)PF * nf;
  const size_t gx_side = (size_t)nr * n * 3, gy_side = (size_t)nr * 3 * n;
  const size_t fst = (size_t)12 * gx_side;
  const bool colw = e < 2;                       // W/E: three edge columns, all rows
  const int c0 = (e == 1) ? n - 3 : 0, r0 = (e == 3) ? n - 3 : 0;
  // A thread past the strip takes point 0's indices: its loads, LDS writes
  // and stores repeat that point's with the same values, so none of them
  // sits under a branch.
  int src[NI], slot[NI];
  size_t ob[NI];
  int ol[NI], om[NI];
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int t = threadIdx.x + it * blockDim.x;
    const int tt = t < strip ? t : 0;
    if (colw) {  // row tt/3, column c0 + tt%3 -> line tt%3, along index tt/3
      src[it] = (tt / 3) * n + c0 + tt % 3;
      slot[it] = (tt % 3) * n + tt / 3;
      ol[it] = tt % 3 + 1;
      om[it] = tt / 3;
      ob[it] = F * fst + (size_t)(P * 2 + e) * gx_side + (size_t)om[it] * 3 + (ol[it] - 1);  // + k * 3n
    } else {     // row r0 + tt/n, column tt%n -> line tt/n, along index tt%n
      src[it] = (r0 + tt / n) * n + tt % n;
      slot[it] = tt;
      ol[it] = tt / n + 1;
      om[it] = tt % n;
      ob[it] = F * fst + (size_t)(P * 2 + (e - 2)) * gy_side + (size_t)(ol[it] - 1) * n + om[it];  // + k * 3n
    }
  }
  double nxt[NI];
  auto load = [&](int k) {
#pragma unroll
    for (int it = 0; it < NI; ++it) nxt[it] = g[(size_t)k * plane + src[it]];
  };
  load(k0);
  // every thread writes a tap, without a branch (read after the first
  // plane's barrier)
  sft[threadIdx.x % 7] = T.filt[threadIdx.x % 7];
  for (int k = k0; k < k1; ++k) {
    double* ln = sl + ((k - k0) & 1) * strip;
#pragma unroll
    for (int it = 0; it < NI; ++it) ln[slot[it]] = nxt[it];
    load(min(k + 1, k1 - 1));  // past the last plane: loaded and not used
    __syncthreads();
    const bool live = k >= kNstr && k <= nr - 1 - kNstr;
#pragma unroll
    for (int it = 0; it < NI; ++it) {
      const int l = ol[it], m = om[it];
      double s = 0.0;
      if (live) {
        if (e == 2) {  // rows 0..2 at column m
          for (int jj = 0; jj <= 3 - l; ++jj) s += sft[3 - l - jj] * ln[jj * n + m];
        } else if (e == 3) {  // rows n-3..n-1 at column m
          for (int jj = n - 4 + l; jj <= n - 1; ++jj) s += sft[n + 2 + l - jj] * ln[(jj - r0) * n + m];
        } else {  // abar on the edge columns at row m
          const int ia = (e == 0) ? 0 : n - 4 + l, ib = (e == 0) ? 3 - l : n - 1;
          for (int ii = ia; ii <= ib; ++ii) {
            double av = 0.0;  // abar(ii, m) = sum_q f[6-q] bbar(ii, m-3+q), zero beyond the panel
            for (int q = 0; q < 7; ++q) {
              const int jj = m - 3 + q;
              if (jj >= 0 && jj < n) av += sft[6 - q] * ln[(ii - c0) * n + jj];
            }
            s += ((e == 0) ? sft[3 - l - ii] : sft[n + 2 + l - ii]) * av;
          }
        }
      }
      (colw ? fgxa : fgya)[ob[it] + (size_t)k * 3 * n] = s;
    }
  }
}

template <int TI, int TJ>
__global__ __launch_bounds__(TI* TJ) void k_adj_filter(DevTables T, const double* __restrict__ bin,
                                                        double* __restrict__ out, int kchunk, int nkch) {
  constexpr int NT = TI * TJ;
  constexpr int NCELL = (TJ + 6) * (TI + 6);
  constexpr int NC = (NCELL + NT - 1) / NT;
  __shared__ double sb[2][TJ + 6][TI + 6];  // alternating by plane: two barriers per plane
  __shared__ double sa[TJ][TI + 6];
  const int n = T.n, nr = T.nr;
  const int nti = n / TI, ntj = n / TJ;
  const int nb = nti * ntj * 30 * nkch;
  int b = xcd_remap(blockIdx.x, nb);
  const int bi = b % nti;
  b /= nti;
  const int bj = b % ntj;
  b /= ntj;
  const int PF = b % 30;
  const int ch = b / 30;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = bi * TI, j0 = bj * TJ;
  const int i = i0 + tx, j = j0 + ty;
  const int k0 = ch * kchunk;
  const int k1 = min(k0 + kchunk, nr);
  if (k0 >= nr) return;
  const size_t plane = (size_t)n * n, nf = plane * nr;
  const double* g = bin + (size_t)PF * nf;
  double* o = out + (size_t)PF * nf;
  const size_t ij = (size_t)j * n + i;

  // cells of the tile this thread loads
  const int tid = ty * TI + tx;
  long long coff[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int e = tid + c * NT;
    coff[c] = -1;
    if (e < NCELL) {
      const int rr = e / (TI + 6), cc = e % (TI + 6);
      const int gj = j0 - 3 + rr, gi = i0 - 3 + cc;
      if (gj >= 0 && gj < n && gi >= 0 && gi < n) coff[c] = (long long)gj * n + gi;
    }
  }
  auto load_plane = [&](int kk, double (&dst)[NC]) {
#pragma unroll
    for (int c = 0; c < NC; ++c) dst[c] = coff[c] >= 0 ? g[(size_t)kk * plane + coff[c]] : 0.0;
  };

  double win[11];
#pragma unroll
  for (int q = 0; q < 11; ++q) win[q] = 0.0;
  // planes kk carrying FILTER_R output, kk in [max(k0-5,5), min(k1+4,nr-6)];
  // the march continues to k1+4 so that output plane k1-1 is formed.
  const int kfirst = max(k0 - kNstr, kNstr), klast = min(k1 + kNstr - 1, nr - 1 - kNstr);
  const int kstop = k1 + kNstr - 1;
  // two planes in flight: nx1 holds plane kk+1, nx2 plane kk+2
  double nx1[NC], nx2[NC];
  if (kfirst <= klast) load_plane(kfirst, nx1);
  if (kfirst + 1 <= klast) load_plane(kfirst + 1, nx2);
  for (int kk = kfirst; kk <= kstop; ++kk) {
    double v = 0.0;
    if (kk <= klast) {  // block-uniform
      const int pb = (kk - kfirst) & 1;
      double (&sbp)[TJ + 6][TI + 6] = sb[pb];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int e = tid + c * NT;
        if (e < NCELL) (&sbp[0][0])[e] = nx1[c];
        nx1[c] = nx2[c];
      }
      if (kk + 2 <= klast) load_plane(kk + 2, nx2);
      __syncthreads();
      // FILTER_ETA^T: abar(row, c) = sum_q f[6-q] bbar(row-3+q, c), in-panel rows
      for (int cc = tx; cc < TI + 6; cc += TI) {
        double a = 0.0;
#pragma unroll
        for (int q = 0; q < 7; ++q) a += T.filt[6 - q] * sbp[ty + q][cc];
        sa[ty][cc] = a;
      }
      __syncthreads();
      // FILTER_XI^T
#pragma unroll
      for (int q = 0; q < 7; ++q) v += T.filt[6 - q] * sa[ty][tx + q];
    }
#pragma unroll
    for (int q = 0; q < 10; ++q) win[q] = win[q + 1];
    win[10] = v;
    // win[q] = A(kk-10+q); output plane kk-5 takes filt_r[10-q] * win[q]
    const int kout = kk - kNstr;
    if (kout >= k0 && kout < k1) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < 11; ++q) s += T.filtr[10 - q] * win[q];
      __builtin_nontemporal_store(s, o + (size_t)kout * plane + ij);
    }
  }
  // kfirst - 5 <= k0 and kstop - 5 = k1 - 1: every plane of [k0, k1) was written
}

// Receiver sampling and adjoint sources.
__global__ void k_sample(const double* __restrict__ q, const long long* __restrict__ idx, int nrec,
                         double* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nrec) out[r] = q[idx[r]];
}
__global__ void k_inject(double* __restrict__ lam, const long long* __restrict__ idx, int nrec,
                         const double* __restrict__ seis, const double* __restrict__ data) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nrec) lam[idx[r]] += 2.0 * (seis[r] - data[r]);
}

// ------------------------------------------------------------------ launchers

static int adj_ti(int n) { return (n % 32 == 0) ? 32 : 16; }

template <int TI, int TJ, bool D32>
static hipError_t launch_adj_stage_t(const DevTables& T, const AdjStageArgs& a, hipStream_t st) {
  const int tiles = (T.n / TI) * (T.n / TJ) * 6;
  // chunks of >= 24 planes, as many as fit in three rounds of the GPU's blocks
  // (two 4-wave blocks per CU, 256 CUs).  Measured at 6x96^2x192 (216 tiles),
  // adjoint step: 7 chunks 3.02 ms, 9 3.08, 8 3.12, 16 3.15, 6 3.19, 5 3.38.
  int nkch = 1536 / tiles;
  const int maxk = T.nr / 24 > 0 ? T.nr / 24 : 1;
  if (nkch > maxk) nkch = maxk;
  if (a.chunks > 0) nkch = std::min(a.chunks, T.nr);
  if (nkch < 1) nkch = 1;
  const int kchunk = (T.nr + nkch - 1) / nkch;
  hipLaunchKernelGGL((k_adj_stage<TI, TJ, D32>), dim3(tiles * nkch), dim3(TI, TJ), 0, st, T, a.W, a.Yin, a.Yout,
                     a.Wout, a.divs, a.gc2, a.c2f, a.gxa_p, a.gya_p, a.gxa_v, a.gya_v, a.dt, a.alpha, a.beta, a.wscale,
                     kchunk, nkch, a.mode);
  return hipGetLastError();
}

hipError_t launch_adj_stage(const DevTables& T, const AdjStageArgs& a, hipStream_t st) {
  const bool d32 = a.div32 && a.divs && (a.mode & kAdjGrad);
  if (adj_ti(T.n) == 32)
    return d32 ? launch_adj_stage_t<32, 8, true>(T, a, st) : launch_adj_stage_t<32, 8, false>(T, a, st);
  return d32 ? launch_adj_stage_t<16, 16, true>(T, a, st) : launch_adj_stage_t<16, 16, false>(T, a, st);
}

template <bool VEC, bool RADF, int NI, bool HASW>
static hipError_t launch_gathert_w(const DevTables& T, const double* gxs, const double* gys, const double* gxv,
                                   const double* gyv, double* Y, double* Wn, double beta, int sfield, int nfields,
                                   hipStream_t st) {
  const long long field_stride = 12LL * T.nr * T.n * 3;
  dim3 grid((unsigned)(12 * gather_blocks_per_edge(T.nr)), (unsigned)nfields);
  const int nt = std::min(gathert_threads_max(VEC, NI), ((3 * T.n + NI - 1) / NI + 63) / 64 * 64);
  const size_t lds = sizeof(double) * 2 * (VEC ? 9 : 3) * (size_t)T.n;  // n <= 512: at most 72 KiB
  if (lds > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_adj_gathert<VEC, RADF, NI, HASW>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
  }
  for (int pass = 0; pass < 2; ++pass) {
    hipLaunchKernelGGL((k_adj_gathert<VEC, RADF, NI, HASW>), grid, dim3(nt), lds, st, T, gxs, gys, gxv, gyv, Y, Wn,
                       beta, sfield, field_stride, pass);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <bool VEC, bool RADF, int NI>
static hipError_t launch_gathert(const DevTables& T, const double* gxs, const double* gys, const double* gxv,
                                 const double* gyv, double* Y, double* Wn, double beta, int sfield, int nfields,
                                 hipStream_t st) {
  if (Wn) return launch_gathert_w<VEC, RADF, NI, true>(T, gxs, gys, gxv, gyv, Y, Wn, beta, sfield, nfields, st);
  return launch_gathert_w<VEC, RADF, NI, false>(T, gxs, gys, gxv, gyv, Y, Wn, beta, sfield, nfields, st);
}

hipError_t launch_adj_scatter(const DevTables& T, bool vec, const double* gxs, const double* gys, const double* gxv,
                              const double* gyv, double* Y, double* Wn, double beta, int sfield, int nfields,
                              hipStream_t st, bool radial_filter) {
  if (vec) {  // strip points per thread: 1 (n <= 341, up to 1024 threads), 3 (n <= 512, 512 threads)
    if (3 * T.n <= 1024)
      return launch_gathert<true, false, 1>(T, gxs, gys, gxv, gyv, Y, Wn, beta, sfield, nfields, st);
    return launch_gathert<true, false, 3>(T, gxs, gys, gxv, gyv, Y, Wn, beta, sfield, nfields, st);
  }
  const bool two = 3 * T.n > 1024;  // n > 341: two strip points per thread
  if (radial_filter)
    return two ? launch_gathert<false, true, 2>(T, gxs, gys, gxv, gyv, Y, Wn, beta, sfield, nfields, st)
               : launch_gathert<false, true, 1>(T, gxs, gys, gxv, gyv, Y, Wn, beta, sfield, nfields, st);
  return two ? launch_gathert<false, false, 2>(T, gxs, gys, gxv, gyv, Y, Wn, beta, sfield, nfields, st)
             : launch_gathert<false, false, 1>(T, gxs, gys, gxv, gyv, Y, Wn, beta, sfield, nfields, st);
}

template <int TI, int TJ>
static hipError_t launch_adj_filter_t(const DevTables& T, const double* in, double* out, double* fgxa, double* fgya,
                                      hipStream_t st, int chunks) {
  const int tiles = (T.n / TI) * (T.n / TJ) * 30;
  // chunks of >= 48 planes: the march re-reads 10 planes per chunk
  int nkch = (2048 + tiles - 1) / tiles;
  const int maxk = T.nr / 48 > 0 ? T.nr / 48 : 1;
  if (nkch > maxk) nkch = maxk;
  if (chunks > 0) nkch = std::min(chunks, T.nr);
  const int kchunk = (T.nr + nkch - 1) / nkch;
  hipLaunchKernelGGL((k_adj_filter<TI, TJ>), dim3(tiles * nkch), dim3(TI, TJ), 0, st, T, in, out, kchunk, nkch);
  const size_t lds = sizeof(double) * 2 * 3 * (size_t)T.n;
  const dim3 grid((unsigned)(120 * gather_blocks_per_edge(T.nr)));
  if (3 * T.n > 1024)
    hipLaunchKernelGGL(k_adj_filter_ghosts<2>, grid, dim3(gather_threads(T.n, 2)), lds, st, T, in, fgxa, fgya);
  else
    hipLaunchKernelGGL(k_adj_filter_ghosts<1>, grid, dim3(gather_threads(T.n, 1)), lds, st, T, in, fgxa, fgya);
  return hipGetLastError();
}

hipError_t launch_adj_filter(const DevTables& T, const double* in, double* out, double* fgxa, double* fgya,
                             hipStream_t st, int chunks) {
  if (T.n % 32 == 0) return launch_adj_filter_t<32, 16>(T, in, out, fgxa, fgya, st, chunks);
  return launch_adj_filter_t<16, 16>(T, in, out, fgxa, fgya, st, chunks);
}

hipError_t launch_sample(const double* q, const long long* idx, int nrec, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_sample, dim3((nrec + 63) / 64), dim3(64), 0, st, q, idx, nrec, out);
  return hipGetLastError();
}

hipError_t launch_inject(double* lam, const long long* idx, int nrec, const double* seis, const double* data,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_inject, dim3((nrec + 63) / 64), dim3(64), 0, st, lam, idx, nrec, seis, data);
  return hipGetLastError();
}

}  // namespace solar
