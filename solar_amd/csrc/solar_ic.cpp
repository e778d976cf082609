// Seeded initial conditions and radial backgrounds.
//
// The reference reads its background from a `solar_model` text file and
// nondimensionalizes it by the values at the first radial point
// (ic_acoustic_wave.f90:60-108), and fills q in GENERATE_IC
// (ic_acoustic_wave.f90:4-29, ic_gaussian_pulse.f90:1-30).  These are the
// caller-side inputs of the hot path; this file produces them from closed
// forms with a counter-based RNG, so the same seed gives the same bits on
// every machine with the same libm.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "solar_tables.hpp"

namespace {

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Uniform in (0,1) from (seed, stream, counter).
double uniform01(uint64_t seed, uint64_t stream, uint64_t counter) {
  const uint64_t h = splitmix64(splitmix64(seed ^ (stream * 0xD1342543DE82EF95ull)) + counter);
  return (static_cast<double>(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

double normal01(uint64_t seed, uint64_t stream, uint64_t counter) {
  const double u1 = uniform01(seed, stream, 2 * counter);
  const double u2 = uniform01(seed, stream, 2 * counter + 1);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * std::acos(-1.0) * u2);
}

// Unit position of grid point (i,j) of panel p: the reference's panel
// spherical coordinates (metric.f90:97-158) correspond to the cube faces
// 1:(1,X,Y) 2:(-X,1,Y) 3:(-1,-X,Y) 4:(X,-1,Y) 5:(-Y,X,1) 6:(Y,X,-1)
// with X = tan(xi), Y = tan(eta).
void cube_point(int p, double X, double Y, double* v) {
  double x, y, z;
  switch (p) {
    case 0: x = 1; y = X; z = Y; break;
    case 1: x = -X; y = 1; z = Y; break;
    case 2: x = -1; y = -X; z = Y; break;
    case 3: x = X; y = -1; z = Y; break;
    case 4: x = -Y; y = X; z = 1; break;
    default: x = Y; y = X; z = -1; break;
  }
  const double nrm = std::sqrt(x * x + y * y + z * z);
  v[0] = x / nrm;
  v[1] = y / nrm;
  v[2] = z / nrm;
}

// Orthonormal real spherical harmonic Y_lm(theta, phi), m in [-l, l], by the
// standard stable recurrence for the normalised associated Legendre functions.
double real_ylm(int l, int m, double theta, double phi) {
  const double pi = std::acos(-1.0);
  const int am = m < 0 ? -m : m;
  const double x = std::cos(theta), s = std::sin(theta);
  // P_m^m normalised
  double pmm = std::sqrt(1.0 / (4.0 * pi));
  for (int k = 1; k <= am; ++k) pmm *= -std::sqrt((2.0 * k + 1.0) / (2.0 * k)) * s;
  double plm;
  if (l == am) {
    plm = pmm;
  } else {
    double pm1 = std::sqrt(2.0 * am + 3.0) * x * pmm;
    if (l == am + 1) {
      plm = pm1;
    } else {
      double pm2 = pmm;
      plm = pm1;
      for (int ll = am + 2; ll <= l; ++ll) {
        const double a = std::sqrt((4.0 * ll * ll - 1.0) / (static_cast<double>(ll * ll - am * am)));
        const double b = std::sqrt(((ll - 1.0) * (ll - 1.0) - am * am) / (4.0 * (ll - 1.0) * (ll - 1.0) - 1.0));
        plm = a * (x * pm1 - b * pm2);
        pm2 = pm1;
        pm1 = plm;
      }
    }
  }
  if (m == 0) return plm;
  const double r2 = std::sqrt(2.0);
  return m > 0 ? r2 * plm * std::cos(am * phi) : r2 * plm * std::sin(am * phi);
}

}  // namespace

extern "C" {

// Radial grid r in [0.7, 1.0] and nondimensional background profiles.
//   kind 0: uniform c = 1, rho0 = 1, g0 = 0.
//   kind 1: stratified, c^2 = (1.05-r)/(1.05-0.7), rho0 = (c^2)^1.5,
//           hydrostatic g0 = 2.5/(0.35*Gamma1) with Gamma1 = 5/3 (the
//           pressure p0 = rho0 c^2/Gamma1 then satisfies dp0/dr = -rho0 g0).
// Units follow ic_acoustic_wave.f90:73-108: radius in units of the outer
// radius, speeds in units of c(r_min), densities in units of rho0(r_min).
int solar_background(int kind, int nr, double* gr, double* c2, double* rho0, double* g0) {
  if (nr < 2 || (kind != 0 && kind != 1)) return -1;
  for (int k = 0; k < nr; ++k) {
    const double r = 0.7 + 0.3 * static_cast<double>(k) / static_cast<double>(nr - 1);
    gr[k] = r;
    if (kind == 0) {
      c2[k] = 1.0;
      rho0[k] = 1.0;
      g0[k] = 0.0;
    } else {
      // divided by the value at the first point, as SOLAR_DATA does
      const double u = (1.05 - r) / (1.05 - 0.7);
      c2[k] = u;
      rho0[k] = std::pow(u, 1.5);
      g0[k] = 2.5 / (0.35 * (5.0 / 3.0));
    }
  }
  return 0;
}

// Unit positions of every grid point, xyz[((p*n + j)*n + i)*3 + c].
int solar_panel_points(int n, double* xyz) {
  if (n < 2) return -1;
  const double pi = std::acos(-1.0);
  for (int p = 0; p < 6; ++p)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const double xi = -pi * 0.25 + ((static_cast<double>(i + 1) - 1.0) * pi * 0.5) / (static_cast<double>(n) - 1.0);
        const double eta = -pi * 0.25 + ((static_cast<double>(j + 1) - 1.0) * pi * 0.5) / (static_cast<double>(n) - 1.0);
        cube_point(p, std::tan(xi), std::tan(eta), &xyz[((static_cast<size_t>(p) * n + j) * n + i) * 3]);
      }
  return 0;
}

// Seeded initial state q(n,n,nr,5,6) (fields rho', v_xi, v_eta, v_r, p').
//   kind 0 (Gaussian pulse): p' = amp*exp(-d^2/w^2), rho' = p'/c^2, v = 0,
//          centre at r = 0.85 and a seed-drawn direction, d the 3-D distance.
//          params[0] = amp, params[1] = width (w <= 0 draws w in [0.02,0.08]).
//   kind 1 (spherical-harmonic mode): p' = amp*Y_lm(theta,phi)*
//          sin(pi*(r-r0)/(r1-r0)) + N(0, noise^2), rho' = 0, v = 0, with
//          1 <= l <= lmax and -l <= m <= l seed-drawn.
//          params[0] = amp, params[1] = noise, params[2] = lmax.
// info (optional, 4 doubles): kind 0 -> theta_s, phi_s, width, r_s;
// kind 1 -> l, m, 0, 0.
int solar_generate_ic(int kind, int n, int nr, const double* gr, const double* c2, uint64_t seed,
                      const double* params, double* q, double* info) {
  if (n < 2 || nr < 2) return -1;
  const size_t plane = static_cast<size_t>(n) * n, nf = plane * nr;
  std::memset(q, 0, sizeof(double) * nf * 30);
  const double pi = std::acos(-1.0);
  double* xyz = new double[6 * plane * 3];
  solar_panel_points(n, xyz);
  int rc = 0;
  if (kind == 0) {
    const double amp = params ? params[0] : 1e-3;
    double w = params ? params[1] : 0.05;
    const double ths = std::acos(1.0 - 2.0 * uniform01(seed, 1, 0));
    const double phs = 2.0 * pi * uniform01(seed, 1, 1);
    if (w <= 0.0) w = 0.02 + 0.06 * uniform01(seed, 1, 2);
    const double rs = 0.85;
    const double xs[3] = {rs * std::sin(ths) * std::cos(phs), rs * std::sin(ths) * std::sin(phs), rs * std::cos(ths)};
    if (info) { info[0] = ths; info[1] = phs; info[2] = w; info[3] = rs; }
    for (int p = 0; p < 6; ++p)
      for (int k = 0; k < nr; ++k)
        for (size_t c = 0; c < plane; ++c) {
          const double* u = &xyz[(p * plane + c) * 3];
          double d2 = 0.0;
          for (int a = 0; a < 3; ++a) {
            const double dd = gr[k] * u[a] - xs[a];
            d2 += dd * dd;
          }
          const double pv = amp * std::exp(-d2 / (w * w));
          const size_t base = static_cast<size_t>(p) * 5 * nf + static_cast<size_t>(k) * plane + c;
          q[base + 4 * nf] = pv;
          q[base] = pv / c2[k];
        }
  } else if (kind == 1) {
    const double amp = params ? params[0] : 1e-3;
    const double noise = params ? params[1] : 1e-6;
    const int lmax = params ? static_cast<int>(params[2]) : 20;
    const int l = 1 + static_cast<int>(uniform01(seed, 2, 0) * lmax) % lmax;
    const int m = -l + static_cast<int>(uniform01(seed, 2, 1) * (2 * l + 1)) % (2 * l + 1);
    if (info) { info[0] = l; info[1] = m; info[2] = 0; info[3] = 0; }
    const double r0 = gr[0], r1 = gr[nr - 1];
    double* ang = new double[6 * plane];
    for (int p = 0; p < 6; ++p)
      for (size_t c = 0; c < plane; ++c) {
        const double* u = &xyz[(p * plane + c) * 3];
        const double th = std::acos(u[2] > 1.0 ? 1.0 : (u[2] < -1.0 ? -1.0 : u[2]));
        const double ph = std::atan2(u[1], u[0]);
        ang[p * plane + c] = real_ylm(l, m, th, ph);
      }
    for (int p = 0; p < 6; ++p)
      for (int k = 0; k < nr; ++k) {
        const double env = std::sin(pi * (gr[k] - r0) / (r1 - r0));
        for (size_t c = 0; c < plane; ++c) {
          const size_t base = static_cast<size_t>(p) * 5 * nf + static_cast<size_t>(k) * plane + c;
          const uint64_t ctr = (static_cast<uint64_t>(p) * nr + k) * plane + c;
          q[base + 4 * nf] = amp * ang[p * plane + c] * env + noise * normal01(seed, 3, ctr);
        }
      }
    delete[] ang;
  } else {
    rc = -1;
  }
  delete[] xyz;
  return rc;
}

}  // extern "C"
