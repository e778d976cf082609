// On-disk formats of the step-adjacent output of the reference (output.f90),
// for a whole six-panel state q(n,n,nr,5,6) held as one block per panel.
// The reference's MPI layout at px = py = pr = 1 is one rank per panel:
// rank = panel index 0..5, zoneid = rank + 1 (init_mpi.f90:34-37).
//   WRITE_RESTART  output.f90:86-190   Restart/restart.<ddd>.<zoneid>: int32
//                  gsizes(4) = nx, ny, nr, 5; real*8 time; int32 step; then
//                  q(nx,ny,nr,5) real*8 in Fortran order at byte 28 (the
//                  MPI-IO view offset 5*intSize + realSize).  The reference
//                  has no reader; solar_restart_read is its inverse.
//   WRITE_VISUAL   output.f90:192-240  Visual/<ddd>/vis_<rank>: one
//                  unformatted sequential record (int32 length markers) of
//                  (((q(i,j,o_rad,l), i), j), l).
//   WRITE_GRID     output.f90:242-267  Visual/grid/grid_<rank>: (1x, 7i5).
//   WRITE_RADIAL_TRACE output.f90:312-343  trace_<zoneid>, appended:
//                  ((q(1,1,k,l), k = 1..nr), l = 1..5) in (1x, 5(ES13.5)).
// <ddd> is the cue number as digits cue/100, cue%100/10, cue%10 (output.f90:
// 121-127, dirs.c:21).  Paths are relative to `dir`; missing directories are
// created.
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/solar_amd.h"
#include "solar_internal.hpp"

namespace {

std::string cue_digits(int cue) {
  char b[64];
  std::snprintf(b, sizeof b, "%d%d%d", cue / 100, (cue % 100) / 10, cue % 10);
  return b;
}

// mkdir -p of every component of path (a directory).
bool make_dirs(const std::string& path) {
  std::string cur;
  size_t pos = 0;
  while (pos <= path.size()) {
    const size_t nx = path.find('/', pos);
    cur = path.substr(0, nx == std::string::npos ? path.size() : nx);
    if (!cur.empty() && ::mkdir(cur.c_str(), S_IRWXU | S_IRWXG) != 0 && errno != EEXIST) return false;
    if (nx == std::string::npos) break;
    pos = nx + 1;
  }
  return true;
}

std::string join(const char* dir, const std::string& rel) {
  std::string d = (dir && dir[0]) ? dir : ".";
  if (d.back() != '/') d += '/';
  return d + rel;
}

// Fortran ESw.d edit descriptor with w = 13, d = 5 as gfortran writes it:
// mantissa d.ddddd, exponent E+ee (or +eee without the E beyond 99).
std::string es13_5(double v) {
  char b[64];
  if (std::isnan(v)) {
    std::snprintf(b, sizeof b, "%13s", "NaN");
    return b;
  }
  if (std::isinf(v)) {
    std::snprintf(b, sizeof b, "%13s", v > 0 ? "Infinity" : "-Infinity");
    return b;
  }
  char m[64];
  std::snprintf(m, sizeof m, "%.5E", v);  // [-]d.dddddE[+-]dd[d]
  const char* e = std::strchr(m, 'E');
  const std::string mant(m, e - m);
  const int ex = std::atoi(e + 1);
  char ebuf[16];
  if (std::abs(ex) <= 99)
    std::snprintf(ebuf, sizeof ebuf, "E%c%02d", ex < 0 ? '-' : '+', std::abs(ex));
  else
    std::snprintf(ebuf, sizeof ebuf, "%c%03d", ex < 0 ? '-' : '+', std::abs(ex));
  std::snprintf(b, sizeof b, "%13s", (mant + ebuf).c_str());
  return b;
}

int bad(const std::string& m) { return solar::set_error(SOLAR_ERR_ARG, m); }

}  // namespace

extern "C" {

int solar_restart_write(const char* dir, int n, int nr, const double* q, int cue, double time, int step) {
  if (n < 1 || nr < 1 || !q || cue < 0) return bad("solar_restart_write: invalid argument");
  const std::string sub = join(dir, "Restart");
  if (!make_dirs(sub)) return bad("solar_restart_write: cannot create " + sub);
  const size_t block = static_cast<size_t>(n) * n * nr * 5;
  for (int p = 0; p < 6; ++p) {
    const std::string f = sub + "/restart." + cue_digits(cue) + "." + std::to_string(p + 1);
    FILE* fp = std::fopen(f.c_str(), "wb");
    if (!fp) return bad("solar_restart_write: cannot open " + f);
    const int32_t gs[4] = {n, n, nr, 5};
    const int32_t st = step;
    bool ok = std::fwrite(gs, sizeof gs, 1, fp) == 1 && std::fwrite(&time, sizeof time, 1, fp) == 1 &&
              std::fwrite(&st, sizeof st, 1, fp) == 1 &&
              std::fwrite(q + p * block, sizeof(double), block, fp) == block;
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) return bad("solar_restart_write: write failed on " + f);
  }
  return SOLAR_OK;
}

int solar_restart_read(const char* dir, int n, int nr, double* q, int cue, double* time, int* step) {
  if (n < 1 || nr < 1 || !q || cue < 0) return bad("solar_restart_read: invalid argument");
  const size_t block = static_cast<size_t>(n) * n * nr * 5;
  double t0 = 0.0;
  int32_t s0 = 0;
  for (int p = 0; p < 6; ++p) {
    const std::string f = join(dir, "Restart/restart." + cue_digits(cue) + "." + std::to_string(p + 1));
    FILE* fp = std::fopen(f.c_str(), "rb");
    if (!fp) return bad("solar_restart_read: cannot open " + f);
    int32_t gs[4] = {0, 0, 0, 0}, st = 0;
    double t = 0.0;
    bool ok = std::fread(gs, sizeof gs, 1, fp) == 1 && std::fread(&t, sizeof t, 1, fp) == 1 &&
              std::fread(&st, sizeof st, 1, fp) == 1;
    if (ok && (gs[0] != n || gs[1] != n || gs[2] != nr || gs[3] != 5)) {
      std::fclose(fp);
      return bad("solar_restart_read: " + f + " holds a " + std::to_string(gs[0]) + "x" + std::to_string(gs[1]) +
                 "x" + std::to_string(gs[2]) + "x" + std::to_string(gs[3]) + " array, expected " +
                 std::to_string(n) + "x" + std::to_string(n) + "x" + std::to_string(nr) + "x5");
    }
    ok = ok && std::fread(q + p * block, sizeof(double), block, fp) == block;
    std::fclose(fp);
    if (!ok) return bad("solar_restart_read: short file " + f);
    if (p == 0) {
      t0 = t;
      s0 = st;
    } else if (t != t0 || st != s0) {
      return bad("solar_restart_read: time/step differ between zone files of cue " + std::to_string(cue));
    }
  }
  if (time) *time = t0;
  if (step) *step = s0;
  return SOLAR_OK;
}

int solar_visual_write(const char* dir, int n, int nr, const double* q, int cue, int o_rad) {
  if (n < 1 || nr < 1 || !q || cue < 0) return bad("solar_visual_write: invalid argument");
  if (o_rad < 1 || o_rad > nr) return bad("solar_visual_write: o_rad outside 1..nr");
  const std::string sub = join(dir, "Visual/" + cue_digits(cue));
  if (!make_dirs(sub)) return bad("solar_visual_write: cannot create " + sub);
  const size_t plane = static_cast<size_t>(n) * n, block = plane * nr * 5;
  std::vector<double> rec(plane * 5);
  for (int p = 0; p < 6; ++p) {
    for (int l = 0; l < 5; ++l)
      std::memcpy(&rec[l * plane], q + p * block + (static_cast<size_t>(l) * nr + (o_rad - 1)) * plane,
                  plane * sizeof(double));
    const std::string f = sub + "/vis_" + std::to_string(p);
    FILE* fp = std::fopen(f.c_str(), "wb");
    if (!fp) return bad("solar_visual_write: cannot open " + f);
    const int32_t len = static_cast<int32_t>(rec.size() * sizeof(double));
    bool ok = std::fwrite(&len, sizeof len, 1, fp) == 1 &&
              std::fwrite(rec.data(), sizeof(double), rec.size(), fp) == rec.size() &&
              std::fwrite(&len, sizeof len, 1, fp) == 1;
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) return bad("solar_visual_write: write failed on " + f);
  }
  return SOLAR_OK;
}

int solar_grid_write(const char* dir, int n, int nr) {
  if (n < 1 || nr < 1) return bad("solar_grid_write: invalid argument");
  const std::string sub = join(dir, "Visual/grid");
  if (!make_dirs(sub)) return bad("solar_grid_write: cannot create " + sub);
  for (int p = 0; p < 6; ++p) {
    const std::string f = sub + "/grid_" + std::to_string(p);
    FILE* fp = std::fopen(f.c_str(), "w");
    if (!fp) return bad("solar_grid_write: cannot open " + f);
    std::fprintf(fp, " %5d%5d%5d%5d%5d%5d%5d\n", 0, n - 1, 0, n - 1, 0, nr - 1, p + 1);
    if (std::fclose(fp) != 0) return bad("solar_grid_write: write failed on " + f);
  }
  return SOLAR_OK;
}

int solar_radial_trace_write(const char* dir, int n, int nr, const double* q) {
  if (n < 1 || nr < 1 || !q) return bad("solar_radial_trace_write: invalid argument");
  const size_t plane = static_cast<size_t>(n) * n, block = plane * nr * 5;
  for (int p = 0; p < 6; ++p) {
    const std::string f = join(dir, "trace_" + std::to_string(p + 1));
    FILE* fp = std::fopen(f.c_str(), "a");
    if (!fp) return bad("solar_radial_trace_write: cannot open " + f);
    std::string out = " ";  // the 1x of the first record only: reversion restarts at 5(ES13.5)
    int on_line = 0;
    for (int l = 0; l < 5; ++l)
      for (int k = 0; k < nr; ++k) {
        if (on_line == 5) {
          out += '\n';
          on_line = 0;
        }
        out += es13_5(q[p * block + (static_cast<size_t>(l) * nr + k) * plane]);  // q(1,1,k,l)
        ++on_line;
      }
    out += '\n';
    const bool ok = std::fputs(out.c_str(), fp) >= 0;
    if (std::fclose(fp) != 0 || !ok) return bad("solar_radial_trace_write: write failed on " + f);
  }
  return SOLAR_OK;
}

int solar_output_radius(int nr, const double* gr, double obsheight, int* o_rad) {
  if (nr < 1 || !gr || !o_rad) return bad("solar_output_radius: invalid argument");
  // ic_acoustic_wave.f90:128-133: the last k with the smallest |gr(k) - obsheight| (ties go to larger k)
  double temp = 1.0;
  int k_out = 0;
  for (int k = 1; k <= nr; ++k)
    if (std::fabs(gr[k - 1] - obsheight) <= temp) {
      temp = std::fabs(gr[k - 1] - obsheight);
      k_out = k;
    }
  if (k_out == 0) return bad("solar_output_radius: no radius within 1 of obsheight");
  *o_rad = k_out;
  return SOLAR_OK;
}

}  // extern "C"
