// Host table builder for the acoustic RK4 hot path.  See solar_tables.hpp.
// Compiled with -ffp-contract=off: the expression order below is the
// reference's, so the tables match its in-loop values bit for bit.
#include "solar_tables.hpp"

#include <algorithm>
#include <cmath>

namespace solar {

// header:40-46 (Tam & Webb DRP).
const double kCoeffs1[7] = {-0.02651995, 0.18941314, -0.79926643, 0.0, 0.79926643, -0.18941314, 0.02651995};
// header:49-55, bdy_coeffs1(nst,nstencil) filled column-major: [row k][tap].
const double kBdyCoeffs1[3][7] = {
    {-2.192280339, 4.748611401, -5.108851915, 4.461567104, -2.833498741, 1.128328861, -0.203876371},
    {-0.209337622, -1.084875676, 2.147776050, -1.388928322, 0.768949766, -0.28181465, 0.048230454},
    {0.049041958, -0.468840357, -0.474760914, 1.273274737, -0.518484526, 0.166138533, -0.026369431}};
// header:68-76 (2N-storage RK(5,4)).
const double kAlphaRk4[5] = {0.0, -4.801594388477952e-01, -1.404247195200064e+00, -2.016477077503357e+00,
                             -1.056444269767207e+00};
const double kBetaRk4[5] = {1.028639988104959e-01, 7.408540575766301e-01, 7.426530946683799e-01,
                            4.694937902357905e-01, 1.881733382887932e-01};
// header:99-105 and header:123-133.
const double kFiltCoeffs[7] = {0.015625, -0.09375, 0.234375, 0.6875, 0.234375, -0.09375, 0.015625};
const double kFiltCoeffsR[11] = {0.0009765, -0.0097655, 0.043945, -0.1171875, 0.2050775, 0.753906,
                                 0.2050775, -0.1171875, 0.043945, -0.0097655, 0.0009765};

double sponge_sigma(int gk, int nr, int sponge_len, double a_sponge) {
  // sponge.f90:27-46: REAL() ratios are single precision, the cube is taken
  // in double precision with a real exponent.
  const int lower = sponge_len, upper = nr - sponge_len + 1;
  volatile float num, den;
  if (gk >= upper) {
    num = static_cast<float>(gk - upper);
  } else if (gk <= lower) {
    num = static_cast<float>(lower - gk);
  } else {
    return 0.0;
  }
  den = static_cast<float>(sponge_len);
  const float ratio = num / den;
  return a_sponge * std::pow(static_cast<double>(ratio), 3.0);
}

namespace {

// Panel connectivity.  Derived from the intercommunicator send/receive pairs
// of DDX/DDY (derivative.f90:55-133,374-452) with px=py=1, the receiver-side
// reversals (derivative.f90:217-282,538-561) and the strip arguments each
// call passes (grad.f90:45-46,121-122).  Sender strips: W = columns 2..4,
// E = columns N-1..N-3, S = rows 2..4, N = rows N-1..N-3.
const Route kRoutes[6][4] = {
    // receiver:      W              E              S              N
    /* 1 */ {{3, kE, 0}, {1, kW, 0}, {5, kN, 0}, {4, kS, 0}},
    /* 2 */ {{0, kE, 0}, {2, kW, 0}, {5, kE, 1}, {4, kE, 0}},
    /* 3 */ {{1, kE, 0}, {3, kW, 0}, {5, kS, 1}, {4, kN, 1}},
    /* 4 */ {{2, kE, 0}, {0, kW, 0}, {5, kW, 0}, {4, kW, 1}},
    /* 5 */ {{3, kN, 1}, {1, kN, 0}, {0, kN, 0}, {2, kN, 1}},
    /* 6 */ {{3, kS, 0}, {1, kS, 1}, {2, kS, 1}, {0, kS, 0}},
};

}  // namespace

bool build_tables(int n, int nr, const double* gr, const double* c2, const double* rho0, const double* g0,
                  int sponge_len, double a_sponge, Tables* t, std::string* err) {
  if (n < 16 || n > 512) {
    if (err) *err = "panel size n must lie in [16, 512] (got " + std::to_string(n) + ")";
    return false;
  }
  if (nr < 12 && nr != 1) {  // nr == 1: the reference's 2-D case
    if (err) *err = "radial size nr must be 1 or at least 12 (got " + std::to_string(nr) + ")";
    return false;
  }
  const bool twod = nr == 1;
  if (!twod && (sponge_len < 1 || 2 * sponge_len > nr)) {
    if (err) *err = "sponge length must lie in [1, nr/2]";
    return false;
  }
  *t = Tables();
  t->n = n;
  t->nr = nr;
  const double pi = std::acos(-1.0);
  t->pi = pi;
  const size_t n2 = static_cast<size_t>(n) * n;

  // Angular grids, metric.f90:14-17 (equiangular, both edges included).
  t->gxi.resize(n);
  t->geta.resize(n);
  for (int i = 1; i <= n; ++i) {
    const double v = -pi * 0.25 + ((static_cast<double>(i) - 1.0) * pi * 0.5) / (static_cast<double>(n) - 1.0);
    t->gxi[i - 1] = v;
    t->geta[i - 1] = v;
  }
  // Radial coordinate chi, metric.f90:18-21.
  t->gchi.resize(nr);
  for (int i = 1; i <= nr; ++i)
    t->gchi[i - 1] = 0.0 + ((static_cast<double>(i) - 1.0) * 1.0) / (static_cast<double>(nr) - 1.0);
  t->gr.assign(gr, gr + nr);
  if (twod) {
    // metric.f90:35-38 nullifies the r effect in 2-D (r = 1); the radial
    // metric (0/0 and chi(2) there) is unused, as DDR returns 0
    // (derivative.f90:620-623), and is left at 0.
    t->gr[0] = 1.0;
    t->gchi[0] = 0.0;
  }
  t->tanxi.resize(n);
  t->taneta.resize(n);
  for (int i = 0; i < n; ++i) {
    t->tanxi[i] = std::tan(t->gxi[i]);
    t->taneta[i] = std::tan(t->geta[i]);
  }
  t->dxi = t->gxi[1] - t->gxi[0];
  t->deta = t->dxi;
  t->dchi = twod ? 0.0 : t->gchi[1] - t->gchi[0];

  // drdchi: the DRP derivative of r(chi), metric.f90:48-67.
  t->drdchi.assign(nr, 0.0);
  const double inv_dchi = twod ? 0.0 : 1.0 / t->dchi;
  for (int k = 0; k < nr && !twod; ++k) {
    double s = 0.0;
    if (k >= kNst && k < nr - kNst) {
      for (int q = 0; q < 7; ++q) s = s + (inv_dchi * kCoeffs1[q]) * gr[k - kNst + q];
    } else if (k < kNst) {
      for (int q = 0; q < 7; ++q) s = s + (inv_dchi * kBdyCoeffs1[k][q]) * gr[q];
    } else {
      for (int q = 0; q < 7; ++q) s = s - (inv_dchi * kBdyCoeffs1[nr - 1 - k][6 - q]) * gr[nr - 7 + q];
    }
    t->drdchi[k] = s;
  }

  // Angular metric terms, metric.f90:74-85.
  t->metric_c.resize(n);
  t->metric_d.resize(n);
  for (int i = 0; i < n; ++i) {
    t->metric_c[i] = std::pow(1.0 + t->tanxi[i] * t->tanxi[i], 0.5);
    t->metric_d[i] = std::pow(1.0 + t->taneta[i] * t->taneta[i], 0.5);
  }
  t->mdelta.resize(n2);
  t->sqrtd.resize(n2);
  t->xy.resize(n2);
  t->cd.resize(n2);
  t->dc.resize(n2);
  t->tdx.resize(n2);
  t->tde.resize(n2);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const size_t ij = static_cast<size_t>(j) * n + i;
      const double d = (1.0 + t->tanxi[i] * t->tanxi[i]) + t->taneta[j] * t->taneta[j];
      t->mdelta[ij] = d;
      t->sqrtd[ij] = std::pow(d, 0.5);
      t->xy[ij] = t->tanxi[i] * t->taneta[j];
      t->cd[ij] = t->metric_c[i] * t->metric_d[j];
      // grad.f90:128-129 evaluates these quotients inside the loop.
      t->dc[ij] = t->mdelta[ij] / t->cd[ij];
      t->tdx[ij] = std::tan(t->gxi[i]) / t->metric_d[j];
      t->tde[ij] = std::tan(t->geta[j]) / t->metric_c[i];
    }

  // Ghost targets and Lagrange weights: metric.f90:89-91, interpolation.f90:9-54.
  t->interpoints.resize(3 * n);
  for (int l = 1; l <= 3; ++l)
    for (int i = 0; i < n; ++i)
      t->interpoints[(l - 1) * n + i] = std::atan(std::tan(t->geta[i]) / std::tan(t->gxi[n - 1] + l * t->dxi));
  t->interpolant.assign(static_cast<size_t>(27) * n, 0.0);
  auto wref = [&](int it, int i1, int l1) -> double& {
    return t->interpolant[(static_cast<size_t>(l1 - 1) * n + (i1 - 1)) * 9 + (it + 4)];
  };
  for (int l = 1; l <= 3; ++l) {
    wref(l, 1, l) = 1.0;
    wref(-l, n, l) = 1.0;
    for (int i = 2; i <= n - 1; ++i) {
      const int lo = (i < kNord + 1) ? -i + 1 : -kNord;
      const int hi = (i > n - kNord) ? n - i : kNord;
      for (int it = lo; it <= hi; ++it) wref(it, i, l) = 1.0;
      const double x = t->interpoints[(l - 1) * n + (i - 1)];
      for (int ip = lo; ip <= hi; ++ip)
        for (int it = lo; it <= hi; ++it)
          if (it != ip)
            wref(ip, i, l) =
                ((x - t->geta[i + it - 1]) * wref(ip, i, l)) / (t->geta[i + ip - 1] - t->geta[i + it - 1]);
    }
  }
  // Gather weights: interior and truncated stencils as above (their taps
  // outside [lo,hi] are 0); the two end points copy one source point,
  // interpolation.f90:304-305 (func_int(1,l)=input(l+1,l),
  // func_int(N,l)=input(N-l,l)).
  t->gather_w = t->interpolant;
  for (int l = 1; l <= 3; ++l) {
    double* w0 = &t->gather_w[(static_cast<size_t>(l - 1) * n + 0) * 9];
    double* wn = &t->gather_w[(static_cast<size_t>(l - 1) * n + (n - 1)) * 9];
    for (int q = 0; q < 9; ++q) w0[q] = wn[q] = 0.0;
    w0[4 + l] = 1.0;  // along index m=0 copies along index l (0-based)
    wn[4 - l] = 1.0;  // along index m=n-1 copies along index n-1-l
  }

  // transform.f90:142-155.
  for (int l = 1; l <= 3; ++l) {
    t->xm[l] = std::tan(t->gxi[0] - l * t->dxi);
    t->ym[l] = std::tan(t->geta[0] - l * t->deta);
    t->cm[l] = std::sqrt(1.0 + t->xm[l] * t->xm[l]);
    t->dm[l] = std::sqrt(1.0 + t->ym[l] * t->ym[l]);
    t->xp[l] = std::tan(t->gxi[n - 1] + l * t->dxi);
    t->yp[l] = std::tan(t->geta[n - 1] + l * t->deta);
    t->cp[l] = std::sqrt(1.0 + t->xp[l] * t->xp[l]);
    t->dp[l] = std::sqrt(1.0 + t->yp[l] * t->yp[l]);
  }
  t->global_c.resize(n);
  t->global_d.resize(n);
  for (int i = 0; i < n; ++i) {
    const double tx = std::tan(t->gxi[i]), te = std::tan(t->geta[i]);
    t->global_c[i] = std::pow(1.0 + tx * tx, 0.5);
    t->global_d[i] = std::pow(1.0 + te * te, 0.5);
  }

  // (xi,eta) -> (theta,phi) conversion, transform.f90:19-56.
  t->conv.assign(3 * n2 * 6, 0.0);
  for (int type = 0; type < 3; ++type)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const size_t ij = static_cast<size_t>(j) * n + i;
        double* c = &t->conv[(type * n2 + ij) * 6];
        if (type == 0) {
          c[0] = 1.0;
          c[1] = t->xy[ij] / t->cd[ij];
          c[2] = -1.0;
          c[3] = 1.0;
          c[4] = t->sqrtd[ij] / t->cd[ij];
          c[5] = 0.0;
        } else {
          const double factor = 1.0 / std::pow(t->mdelta[ij] - 1.0, 0.5);
          const double a1 = t->tanxi[i] / t->metric_d[j];
          const double a2 = t->taneta[j] / t->metric_c[i];
          const double b1 = (t->sqrtd[ij] * t->taneta[j]) / t->metric_d[j];
          const double b2 = (t->sqrtd[ij] * t->tanxi[i]) / t->metric_c[i];
          c[1] = a1;
          c[2] = a2;
          c[3] = factor;
          if (type == 1) {  // panel 5
            c[0] = factor;
            c[4] = -b1;
            c[5] = b2;
          } else {  // panel 6
            c[0] = -factor;
            c[4] = b1;
            c[5] = -b2;
          }
        }
      }

  // Receiver-basis transform coefficients, transform.f90:161-274.  The
  // metric factors are evaluated at the sender's along-edge index, as the
  // reference does (its receiver reverses the strip afterwards).
  t->trans.assign(static_cast<size_t>(6) * 4 * 3 * n * 2, 0.0);
  for (int s = 0; s < 6; ++s)
    for (int e = 0; e < 4; ++e)
      for (int l = 1; l <= 3; ++l)
        for (int m = 0; m < n; ++m) {
          double* ab = &t->trans[((((static_cast<size_t>(s) * 4 + e) * 3 + (l - 1)) * n) + m) * 2];
          const double tx = std::tan(t->gxi[m]);
          const double te = std::tan(t->geta[m]);
          double a = 0.0, b = 0.0;
          if (s <= 3 && e == kE) {  // east_xi, to the next equatorial panel's W ghost
            b = (t->cm[l] * t->global_d[m]) / std::sqrt((1.0 + t->xm[l] * t->xm[l]) + te * te);
          } else if (s <= 3 && e == kW) {  // west_xi
            b = (t->cp[l] * t->global_d[m]) / std::sqrt((1.0 + t->xp[l] * t->xp[l]) + te * te);
          } else if (s >= 4) {  // polar panels to equatorial N (5) or S (6) ghosts
            const double yv = (s == 4) ? t->yp[l] : t->ym[l];
            a = -1.0;
            b = (tx * yv) / std::sqrt((1.0 + tx * tx) + yv * yv);
          } else if (s == 0 || s == 2) {  // panels 1, 3: eta ghosts of 5 (N strip) / 6 (S strip)
            const double yv = (e == kN) ? ((s == 0) ? t->ym[l] : t->yp[l]) : ((s == 0) ? t->yp[l] : t->ym[l]);
            const double A = std::sqrt(tx * tx + yv * yv);
            const double B = std::sqrt((1.0 + tx * tx) + yv * yv);
            const double pa = (t->metric_c[m] * yv) / A;
            const double pb = ((t->metric_c[m] / A) * tx) / B;
            a = (e == kN) ? pa : -pa;
            b = (e == kN) ? pb : -pb;
          } else {  // panels 2, 4: xi ghosts of 5 (N strip) / 6 (S strip)
            const double xv = (s == 1) ? t->xp[l] : t->xm[l];
            const double A = std::sqrt(xv * xv + te * te);
            const double B = std::sqrt((1.0 + xv * xv) + te * te);
            const double pa = (t->global_d[m] * xv) / A;
            const double pb = ((t->global_d[m] / A) * te) / B;
            a = (e == kN) ? pa : -pa;
            b = (e == kN) ? -pb : pb;
          }
          ab[0] = a;
          ab[1] = b;
        }

  // Radial rows: background, DDR weights, sponge.
  t->radial.assign(nr, RadialRow{});
  std::vector<double> wk(static_cast<size_t>(nr) * 7, 0.0);  // 2-D: DDR weights 0
  for (int k = 0; k < nr && !twod; ++k) {
    double* w = &wk[static_cast<size_t>(k) * 7];
    if (k >= kNst && k < nr - kNst) {
      const double fac2 = inv_dchi / t->drdchi[k];
      for (int q = 0; q < 7; ++q) w[q] = fac2 * kCoeffs1[q];
    } else if (k < kNst) {
      const double fac2 = inv_dchi / t->drdchi[k];
      for (int q = 0; q < 7; ++q) w[q] = fac2 * kBdyCoeffs1[k][q];
    } else {
      const double fac2 = -inv_dchi / t->drdchi[k];
      for (int q = 0; q < 7; ++q) w[q] = fac2 * kBdyCoeffs1[nr - 1 - k][6 - q];
    }
  }
  for (int k = 0; k < nr; ++k) {
    RadialRow& row = t->radial[k];
    row.inv_r = 1.0 / t->gr[k];
    row.r = t->gr[k];
    row.rho0 = rho0[k];
    row.c2rho0 = c2[k] * rho0[k];
    row.g0 = g0[k];
    row.rho0g0 = rho0[k] * g0[k];
    double s = 0.0;  // DDR(rho0), compute_rhs_acoustic.f90:51; 0 in 2-D
    if (!twod) {
      const int base = (k < kNst) ? 0 : (k >= nr - kNst ? nr - 7 : k - kNst);
      for (int q = 0; q < 7; ++q) s = s + wk[static_cast<size_t>(k) * 7 + q] * rho0[base + q];
    }
    row.drho0dr = s;
    row.sigma = twod ? 0.0 : sponge_sigma(k + 1, nr, sponge_len, a_sponge);  // sponge.f90:18
    for (int q = 0; q < 7; ++q) row.w[q] = wk[static_cast<size_t>(k) * 7 + q];
    row.inv_rho0 = 1.0 / row.rho0;
  }
  for (int q = 0; q < 7; ++q) {
    t->fcx[q] = (1.0 / t->dxi) * kCoeffs1[q];
    t->fcy[q] = (1.0 / t->deta) * kCoeffs1[q];
  }
  for (int p = 0; p < 6; ++p)
    for (int e = 0; e < 4; ++e) t->route[p][e] = kRoutes[p][e];
  for (int p = 0; p < 6; ++p)
    for (int e = 0; e < 4; ++e) {
      const Route& r = kRoutes[p][e];
      t->inv_route[r.sender][r.sender_edge] = (p * 4 + e) | (r.reverse << 8);
    }
  t->adj_metric.resize(static_cast<size_t>(n) * n * 4);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const size_t ij = static_cast<size_t>(j) * n + i;
      t->adj_metric[ij * 4 + 0] = t->dc[ij] / t->metric_c[i];
      t->adj_metric[ij * 4 + 1] = t->dc[ij] / t->metric_d[j];
      t->adj_metric[ij * 4 + 2] = t->xy[ij] / t->metric_c[i];
      t->adj_metric[ij * 4 + 3] = t->xy[ij] / t->metric_d[j];
    }
  // Transposed gather rows (taps clamped to the edge like the gather; equal
  // (a, mp) pairs merged; unused entries keep weight 0 at mp = a).
  t->gt_mp.assign(static_cast<size_t>(3) * n * kGtWidth, 0);
  t->gt_w.assign(static_cast<size_t>(3) * n * kGtWidth, 0.0);
  for (int l = 1; l <= 3; ++l)
    for (int a = 0; a < n; ++a) {
      const size_t row = (static_cast<size_t>(l - 1) * n + a) * kGtWidth;
      int cnt = 0;
      for (int q = 0; q < kGtWidth; ++q) t->gt_mp[row + q] = a;
      for (int mp = std::max(0, a - 8); mp <= std::min(n - 1, a + 8); ++mp) {
        double w = 0.0;
        bool any = false;
        for (int tp = 0; tp < 9; ++tp) {
          const int src = std::min(std::max(mp - 4 + tp, 0), n - 1);
          const double wt = t->gather_w[(static_cast<size_t>(l - 1) * n + mp) * 9 + tp];
          if (src == a && wt != 0.0) {
            w += wt;
            any = true;
          }
        }
        if (any) {
          if (cnt == kGtWidth) {
            if (err) *err = "transposed gather row wider than kGtWidth";
            return false;
          }
          t->gt_mp[row + cnt] = mp;
          t->gt_w[row + cnt] = w;
          ++cnt;
        }
      }
    }
  return true;
}

}  // namespace solar
