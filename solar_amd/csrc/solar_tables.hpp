// Host-side, time-invariant tables of the acoustic RK4 hot path.
//
// Everything the reference re-evaluates inside its loops with libm (tan, sqrt,
// pow, atan) or with per-call divisions of constants is evaluated once here,
// in the same expression order as the reference, so that the device kernels
// only multiply, add and divide by tabulated values and reproduce the
// reference's rounding exactly.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace solar {

constexpr int kNst = 3;   // stencil half width (header:22-25)
constexpr int kNord = 4;  // Lagrange half width (header:24)
constexpr int kNstr = 5;  // radial filter half width (header:26)

// Panel edges as seen from the panel itself.
enum Edge : int { kW = 0, kE = 1, kS = 2, kN = 3 };

// One directed panel relation: the receiver's ghost strip on one edge comes
// from the sender panel's strip along `sender_edge`, with the along-edge
// index reversed when `reverse` is set.
struct Route {
  int sender;       // 0-based panel
  int sender_edge;  // Edge
  int reverse;
};

// Per-radial-plane constants, packed so a kernel reads one 128-byte record.
struct RadialRow {
  double inv_r;    // 1/r(k)                       grad.f90:52, :128
  double r;        // r(k)                         grad.f90:130
  double rho0;     //                              compute_rhs_acoustic.f90:41
  double c2rho0;   // c2*rho0                      :134
  double g0;       //                              :83
  double rho0g0;   // rho0*g0                      :146
  double drho0dr;  // DDR(rho0)                    :51
  double sigma;    // sponge coefficient           sponge.f90:40-46
  double w[7];     // DDR weights on planes base..base+6 (derivative.f90:656-754)
  double inv_rho0;  // 1/rho0(k), correctly rounded: the divisor of the stage's quotients by rho0
};
static_assert(sizeof(RadialRow) == 16 * sizeof(double), "RadialRow layout");

struct Tables {
  int n = 0, nr = 0;
  double pi = 0, dxi = 0, deta = 0, dchi = 0;
  // 1-D angular tables (length n)
  std::vector<double> gxi, geta, tanxi, taneta, metric_c, metric_d, global_c, global_d;
  // 1-D radial tables (length nr)
  std::vector<double> gr, gchi, drdchi;
  // 2-D angular tables, (j*n + i)
  std::vector<double> mdelta, sqrtd, xy, cd;
  std::vector<double> dc;   // metric_delta/CD            grad.f90:128
  std::vector<double> tdx;  // tan(xi)/metric_D           grad.f90:129
  std::vector<double> tde;  // tan(eta)/metric_C          grad.f90:129
  std::vector<double> interpoints;  // (l*n + i), metric.f90:89-91
  // Lagrange weights, (l*n + m)*9 + t, t = iter+4 (interpolation.f90:9-54).
  std::vector<double> interpolant;
  // Ghost gather weights: same shape as interpolant, with the corner copies
  // (interpolation.f90:183-184,304-305) written as a single unit weight.
  std::vector<double> gather_w;
  // Transpose of the gather along an edge: for layer l and sender along
  // index a, the receiver stencils (mp = receiver along index before any
  // reversal, weight) that read a, as fixed-width rows of kGtWidth entries
  // ((l-1)*n + a)*kGtWidth + q, padded with zero weights.
  std::vector<int> gt_mp;
  std::vector<double> gt_w;
  // Adjoint metric factors per point, interleaved (j*n + i)*4 + c:
  // metric_delta/CD / C, metric_delta/CD / D, XY / C, XY / D.
  std::vector<double> adj_metric;
  // Ghost coordinates of transform.f90:142-152, index l = 1..3.
  double xm[4]{}, xp[4]{}, ym[4]{}, yp[4]{}, cm[4]{}, cp[4]{}, dm[4]{}, dp[4]{};
  // Basis conversion (xi,eta) -> (theta,phi) per panel type (0: panels 1-4,
  // 1: panel 5, 2: panel 6) and point: 6 coefficients
  //   theta = F_t*(A1*v_xi + A2*v_eta), phi = F_p*(B1*v_xi + B2*v_eta)
  // in the order [F_t, A1, A2, F_p, B1, B2]  (transform.f90:19-56).
  std::vector<double> conv;  // ((type*n*n) + j*n + i)*6 + c
  // Receiver-basis transform of an interpolated (theta,phi) strip: per
  // (sender panel, sender edge, layer l-1, along index m) the pair (a, b) with
  //   ghost = a*theta + b*phi           (transform.f90:161-274)
  std::vector<double> trans;  // (((s*4 + e)*3 + l)*n + m)*2 + c
  std::vector<RadialRow> radial;
  double fcx[7]{}, fcy[7]{};  // (1/dxi)*coeffs1, (1/deta)*coeffs1
  Route route[6][4]{};
  // For sender panel s and sender edge se: (receiver P*4 + e) | reverse << 8.
  int inv_route[6][4]{};
};

constexpr int kGtWidth = 12;

// Scheme constants of header:40-133.
extern const double kCoeffs1[7];
extern const double kBdyCoeffs1[3][7];
extern const double kAlphaRk4[5];
extern const double kBetaRk4[5];
extern const double kFiltCoeffs[7];
extern const double kFiltCoeffsR[11];

// Builds every table for an n x n x nr panel with radial-only background
// profiles gr, c2, rho0, g0 (length nr).  Returns false and sets err on
// invalid sizes.
bool build_tables(int n, int nr, const double* gr, const double* c2, const double* rho0, const double* g0,
                  int sponge_len, double a_sponge, Tables* t, std::string* err);

// Sponge coefficient for 1-based radial index gk (sponge.f90:27-46).
double sponge_sigma(int gk, int nr, int sponge_len, double a_sponge);

}  // namespace solar
