// Device kernels of the acoustic RK4 hot path for gfx950 (MI355X).
//
// State layout in HBM: q(n,n,nr,5,6), i fastest, one n*n*nr block per field
// per panel, stored as S = double (FP64 path) or float (FP32-state variant:
// storage only -- every kernel computes in FP64 with FP64 tables and ghosts).  All arithmetic follows the reference's expression order and is
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

#include <algorithm>
#include <type_traits>

#include "solar_device.hpp"
#include "solar_kernels.hpp"

namespace solar {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// a / b rounded to nearest -- the IEEE quotient, bit for bit -- from
// y = RN(1/b), precomputed where b is fixed: q = RN(a y), then two
// corrections q <- RN(q - r y) with the negated remainder r = b q - a, exact
// by FMA once q is within an ulp of a/b (Markstein: with y within half an ulp
// of 1/b and q within an ulp of a/b, the correction rounds to RN(a/b)).  The
// negated remainder keeps the sign of a zero quotient (-0 / b = -0).  The
// remainders and quotients must stay normal: with the divisors in
// [2^-60, 2^60] (fast, checked on the host) a numerator whose exponent lies
// in [-900, 900], or zero, may take this path; others (subnormal field
// values occur in the tails of Gaussian pulses) take the hardware division
// (div_group), as do all when fast is false.  An infinite numerator gives
// NaN here where the division gives an infinity (both non-finite).  Five
// FP64 operations and an exponent test, against about a dozen with a
// quarter-rate reciprocal.
__device__ __forceinline__ double quot_rcp(double a, double b, double y) {
  const double q0 = a * y;
  const double r0 = __builtin_fma(b, q0, -a);
  const double q1 = __builtin_fma(-r0, y, q0);
  const double r1 = __builtin_fma(b, q1, -a);
  return __builtin_fma(-r1, y, q1);
}
// The numerator test: above 1800 when the exponent lies outside [-900, 900].
__device__ __forceinline__ unsigned quot_exp(double a) { return (unsigned)(__builtin_amdgcn_frexp_exp(a) + 900); }
// A group of quotients with one test (RCP): every lane takes the reciprocal
// sequence; in a wave where some lane's numerators fail the test (or fast
// is false) those lanes redo the group by hardware division.  The common
// path then has no branch per quotient.  Without RCP, hardware division.
// Measured in one call: the FP32-state step at c3 16.23 -> 16.03 ms against
// a guard per quotient; the FP64 stages, latency-bound at two waves per
// SIMD, were no faster at c3 and slower at 6x128^2x256 and 6x96^2x192
// (+0.5%, +1.6%) than with hardware division, and keep it (a guard per
// quotient had been slower there too: it drained the prefetch at the
// loop's back edge).
template <bool RCP, int N>
__device__ __forceinline__ void div_group(const double (&a)[N], const double (&b)[N], const double (&y)[N], bool fast,
                                          double (&q)[N]) {
  if constexpr (!RCP) {
#pragma unroll
    for (int t = 0; t < N; ++t) q[t] = a[t] / b[t];
    return;
  }
  unsigned m = 0;
#pragma unroll
  for (int t = 0; t < N; ++t) {
    q[t] = quot_rcp(a[t], b[t], y[t]);
    m = max(m, quot_exp(a[t]));
  }
  const bool slow = !fast || m > 1800u;
  if (__builtin_expect(__ballot(slow) != 0, 0)) {
    if (slow) {
#pragma unroll
      for (int t = 0; t < N; ++t) q[t] = a[t] / b[t];
    }
  }
}

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

// Ghost strips of one receiver edge (receiver panel P, edge e) over the
// planes [k0, k1), all three layers.  Per plane the sender's three source
// lines (rows/columns 2..4 or N-1..N-3 of grad.f90:28-31) are staged in LDS
// -- for vectors already converted to (theta, phi) with the sender's metric
// (transform.f90:19-56); threads walk the layer index fastest so the three
// adjacent columns of a W/E strip share cache lines.  Each thread then
// interpolates ghost values along the edge with the 9-point Lagrange weights
// (interpolation.f90:140-378; unit weights encode the end-point copies) and,
// for vectors, applies the receiver-basis transform (transform.f90:161-274).
// The conversion, weight and transform tables do not depend on the plane:
// each thread loads its own once into registers.  The next plane's source
// values load while this plane is interpolated, and the staged lines are
// double-buffered in LDS, so one barrier per plane suffices.
//   gx*: [P][side W/E][k][m][l],  gy*: [P][side S/N][k][l][m]
//   ecols: when not null, the compact copies of the sender's edge columns
//     (edge_cols layout, written by the stage that produced q); W/E strips
//     are then read from there, one contiguous run per field and plane,
//     instead of three values from every row of the field.
//   NI: strip points per thread (3n <= NI * blockDim.x)
//   sline: 2 buffers x 3 layers x (3 quantities for vectors, else 1) x n
//   qwr (not null, row strips only): the source is q + betar * qwr, the
//     input of a stage that is rebuilt from q and qw (the edge columns hold
//     it as computed).
template <bool VEC, typename S, int NI>
__device__ __forceinline__ void gather_edge(const DevTables& T, int P, int e, int k0, int k1, int fld,
                                            const S* __restrict__ q, const S* __restrict__ ecols, int sfield,
                                            double* __restrict__ gxs, double* __restrict__ gys,
                                            double* __restrict__ gxv, double* __restrict__ gyv,
                                            long long field_stride_out, double* sline, const S* __restrict__ qwr,
                                            double betar) {
  constexpr int NQ = VEC ? 3 : 1;
  const int n = T.n, nr = T.nr;
  const Route rt = T.route[P * 4 + e];
  const int s = rt.sender, se = rt.sender_edge;
  const size_t plane = (size_t)n * n, nf = plane * nr;
  const int colstrip = (se == kW || se == kE);  // column strips: layer index is the fast memory index
  const bool from_cols = colstrip && ecols != nullptr;
  // source of field f at plane k, strip offset x: src[f * fstr + k * kstr + x]
  const S* src = from_cols ? ecols + (size_t)(s * 2 + (se == kE)) * nr * edge_cols_plane(n) : q + (size_t)s * 5 * nf;
  const size_t fstr = from_cols ? (size_t)3 * n : nf, kstr = from_cols ? edge_cols_plane(n) : plane;
  const auto sp = to_global(src + (size_t)(sfield + fld) * fstr);
  const auto vx = to_global(src + fstr);
  const auto ve = vx + fstr;
  const bool rec = qwr != nullptr && !from_cols;  // block-uniform
  const auto wsrc = to_global((rec ? qwr : q) + (size_t)s * 5 * nf);
  const auto wsp = wsrc + (size_t)(sfield + fld) * nf;
  const auto wvx = wsrc + nf;
  const auto wve = wvx + nf;
  const int ctype = s < 4 ? 0 : s - 3;
  const double* cv = T.conv + (size_t)ctype * plane * 6;
  const int lstride = NQ * n, bstride = 3 * lstride;
  const size_t fo = (size_t)fld * field_stride_out;

  // This thread's staged points (l, a) and its outputs (l, m), with their
  // plane-independent tables.
  int sidx[NI], sslot[NI];
  bool sval[NI], oval[NI];
  double c[NI][6];
  int obase_l[NI], omp[NI];
  size_t obase[NI];
  double w[NI][9], ab[NI][2];
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int t = threadIdx.x + it * blockDim.x;
    sval[it] = t < 3 * n;
    oval[it] = sval[it];
    const int tt = sval[it] ? t : 0;
    const int l = colstrip ? tt % 3 + 1 : tt / n + 1;
    const int a = colstrip ? tt / 3 : tt % n;
    const int pidx = strip_index(se, l, a, n);  // point in the sender's plane
    sidx[it] = from_cols ? tt : pidx;            // edge_cols: (a, l) at a * 3 + l - 1
    sslot[it] = (l - 1) * lstride + a;
    if (VEC) {
#pragma unroll
      for (int x = 0; x < 6; ++x) c[it][x] = cv[(size_t)pidx * 6 + x];
    }
    const int lo = (e < 2) ? tt % 3 + 1 : tt / n + 1;  // output layout: l fastest for x ghosts
    const int m = (e < 2) ? tt / 3 : tt % n;
    const int mp = rt.reverse ? n - 1 - m : m;
    obase_l[it] = (lo - 1) * lstride;
    omp[it] = mp;
#pragma unroll
    for (int tp = 0; tp < 9; ++tp) w[it][tp] = T.gw[((size_t)(lo - 1) * n + mp) * 9 + tp];
    if (VEC) {
      const double* abp = T.trans + ((((size_t)s * 4 + se) * 3 + (lo - 1)) * n + mp) * 2;
      ab[it][0] = abp[0];
      ab[it][1] = abp[1];
    }
    obase[it] = (e < 2) ? fo + ((size_t)(P * 2 + e) * nr * n + m) * 3 + (lo - 1)
                        : fo + ((size_t)(P * 2 + (e - 2)) * nr * 3 + (lo - 1)) * n + m;
  }
  const auto gs = to_global(e < 2 ? gxs : gys);
  const auto gv = to_global(e < 2 ? gxv : gyv);

  // source values of one plane, in the storage type until staged
  S pv[NI], xv[NI], yv[NI], pw[NI], xw[NI], yw[NI];
  auto load = [&](int k) {
    const size_t kp = (size_t)k * kstr;
#pragma unroll
    for (int it = 0; it < NI; ++it) {
      pv[it] = sp[kp + sidx[it]];
      if (VEC) {
        xv[it] = vx[kp + sidx[it]];
        yv[it] = ve[kp + sidx[it]];
      }
      if (rec) {
        pw[it] = wsp[kp + sidx[it]];
        if (VEC) {
          xw[it] = wvx[kp + sidx[it]];
          yw[it] = wve[kp + sidx[it]];
        }
      }
    }
  };
  load(k0);
  for (int k = k0; k < k1; ++k) {
    double* buf = sline + ((k - k0) & 1) * bstride;
#pragma unroll
    for (int it = 0; it < NI; ++it) {
      if (!sval[it]) continue;
      double* ln = buf + sslot[it];
      double pvd = static_cast<double>(pv[it]);
      if (rec) pvd = pvd + betar * static_cast<double>(pw[it]);
      ln[0] = pvd;
      if (VEC) {
        double x = static_cast<double>(xv[it]), y = static_cast<double>(yv[it]);
        if (rec) {
          x = x + betar * static_cast<double>(xw[it]);
          y = y + betar * static_cast<double>(yw[it]);
        }
        ln[n] = c[it][0] * (c[it][1] * x + c[it][2] * y);
        ln[2 * n] = c[it][3] * (c[it][4] * x + c[it][5] * y);
      }
    }
    load(min(k + 1, k1 - 1));  // past the last plane: loaded and not used
    __syncthreads();
#pragma unroll
    for (int it = 0; it < NI; ++it) {
      if (!oval[it]) continue;
      const double* ln = buf + obase_l[it];
      double ps = 0.0, ths = 0.0, phs = 0.0;
#pragma unroll
      for (int tp = 0; tp < 9; ++tp) {
        const int a = clampi(omp[it] - 4 + tp, 0, n - 1);
        const double wt = w[it][tp];
        ps = ps + ln[a] * wt;
        if (VEC) {
          ths = ths + ln[n + a] * wt;
          phs = phs + ln[2 * n + a] * wt;
        }
      }
      const size_t o = obase[it] + (size_t)k * 3 * n;
      gs[o] = ps;
      if (VEC) gv[o] = ab[it][0] * ths + ab[it][1] * phs;
    }
  }
}

// blockIdx.x = (chunk of planes)*24 + (receiver panel P)*4 + (receiver edge e);
// blockIdx.y selects a field offset (the filter's five scalars).
// The FP32-state vector gather at six waves per SIMD (80 registers, a 12-byte
// spill): c3 80 against 97 us.  The FP64 one gained nothing at c3 and lost
// at 6x96^2x192 (28 against 23 us) and keeps the compiler's choice.
template <bool VEC, typename S, int NI>
__global__ __launch_bounds__(gather_threads_max(NI))
__attribute__((amdgpu_waves_per_eu(VEC && NI == 1 && sizeof(S) == 4 ? 6 : 1)))
void k_gather(DevTables T, const S* __restrict__ q,
                                                 const S* __restrict__ ecols, int sfield,
                                                 double* __restrict__ gxs, double* __restrict__ gys,
                                                 double* __restrict__ gxv, double* __restrict__ gyv,
                                                 long long field_stride_out, const S* __restrict__ qwr,
                                                 double betar) {
  extern __shared__ double sline[];
  const int kc = gather_kc(T.nr);
  const int k0 = (blockIdx.x / 24) * kc, edge = blockIdx.x % 24;
  gather_edge<VEC, S, NI>(T, edge / 4, edge % 4, k0, min(k0 + kc, T.nr), blockIdx.y, q, ecols, sfield, gxs, gys,
                          gxv, gyv, field_stride_out, sline, qwr, betar);
}

// Values one thread loads for one radial plane of the stage kernel: its own
// point, the halo cells it fills, and the plane entering the radial window.
// State values stay in the storage type S until used (FP32 storage halves
// the registers of the prefetched plane); halo cells may come from the FP64
// ghost strips and are held as double.
template <typename S>
struct StagePlane {
  S vx, ve, q1;                // v_xi, v_eta, rho' at (i,j,k)
  S w[5];                      // qw at (i,j,k)
  double hpx, hvx;             // i-halo cell (p', v_xi)
  double hpy0, hve0, hpy1, hve1;  // low / high j-halo cells (p', v_eta)
  // FP32 state: halo cells inside the panel, kept as loaded and widened when
  // they go to LDS (a conversion at the load would wait on every load in
  // flight); the double members then hold the ghost-strip cells only
  S fpx, fvx, fpy0, fve0, fpy1, fve1;
  S pn, rn;                    // p', v_r of the plane entering the window
  double ogp, ogv;             // PART: ghosts of an outside thread's own cell (p' and v_xi or v_eta)
  double c2;                   // optional 3-D c^2 at (i,j,k)
  double tdc2, tdiv;           // tangent-linear inputs at (i,j,k)
  // RECON: qw at the halo cells and at the plane entering the window
  S rpx, rvx, rpy0, rve0, rpy1, rve1, wpn, wrn;
};

// One RK stage (or one RHS evaluation when rhs_mode is set), one pass over
// HBM.  A block owns a TI x TJ column tile of one panel and marches through a
// chunk of radial planes; p' and v_r sit in 7-plane register windows for the
// radial stencils, the i/j neighbours of plane k come from a double-buffered
// LDS tile, and plane k+1 is loaded into registers while plane k is computed.
//   KIND 0: first stage (qw taken as 0, written), 1: middle stage (qw read and
//   written), 2: last stage (qw read, not written back).
//   XF (FP64 storage, single grid): the stage schedule that stores q after
//   stages 2, 4 and 5 only.  kStoreNoQ: the stage stores qw (in place) and
//   the edge columns of q, but not q itself.  kStageRecon (KIND 1): the stage
//   input is qin + beta_r * qw_a, rebuilt wherever it is read (own point,
//   halo cells, radial window) with the expression the previous stage
//   rounded when it computed it, so the same bits; q goes to qout and qw to
//   qwo_a, a buffer other than qw_a (other blocks still read qw_a's halo
//   cells).
// Occupancy floor of the stage kernel (waves per SIMD); 0 leaves it to the
// compiler.  2 keeps every variant within 256 registers: left alone, the
// ensemble and tangent-linear middle stages take 254-255 VGPRs plus 2 AGPRs
// and drop to one wave per SIMD (measured: ensemble step 44.3 -> 42.4 ms,
// the single-grid stages unchanged).  The single-grid FP32-state variants
// get 3 (168 registers): the middle stage then spills 20 bytes per lane, and
// still runs faster with the third wave (measured at 6x128^2x256: 0.537 ->
// 0.501 ms per stage, step 7.57e9 -> 7.92e9 grid-point steps/s).
#ifndef SOLAR_STAGE_WAVES
#define SOLAR_STAGE_WAVES 2
#endif
#ifndef SOLAR_STAGE_WAVES_F32
#define SOLAR_STAGE_WAVES_F32 3
#endif
#if SOLAR_STAGE_WAVES > 0
#define SOLAR_STAGE_ATTR \
  __attribute__((amdgpu_waves_per_eu(sizeof(S) == 4 && MODE == 0 ? SOLAR_STAGE_WAVES_F32 : SOLAR_STAGE_WAVES)))
#else
#define SOLAR_STAGE_ATTR
#endif

//   ENS: ensemble launch.  The block walks a list of work items (member,
//   tile, panel, plane range) with a grid-sized stride; tables, buffers and
//   dt come from the member (the state ping-pongs between M.q[cur] and
//   M.q[1-cur]); the single-grid arguments are unused.
//   MODE 2: tangent-linear stage (tl_dc2, tl_div used).
//   ecols_a (single grid, not null; in an ensemble the member's ecols): the
//   stage also stores its output at the edge columns i = 1..3 and n-4..n-2
//   in the edge_cols layout, for the next gather -- p', v_xi, v_eta after
//   stages 1-4 (the vector gather), all five fields after the last stage
//   (the filter gather).
//   PART: the tiles need not divide n.  The last tile of a row or column
//   then has threads past the panel edge ("outside" threads): they store
//   nothing, and those within three points of the edge bring the ghost values
//   of their own cell (E ghosts of p' and v_xi, N ghosts of p' and v_eta)
//   into the LDS tile, where the in-panel threads' stencils read them.  Every
//   other cell an outside thread would load reads a fixed in-bounds address.
template <int TI, int TJ, int KIND, typename S, int MODE = 0, bool PART = false, int XF = 0>
__global__ __launch_bounds__(TI* TJ) SOLAR_STAGE_ATTR void k_stage(DevTables Targ, const S* __restrict__ qin_a,
                                                   S* __restrict__ qout_a, S* __restrict__ qw_a,
                                                   const double* __restrict__ gxp_a, const double* __restrict__ gyp_a,
                                                   const double* __restrict__ gxv_a, const double* __restrict__ gyv_a,
                                                   double alpha, double beta, double dt_a, int kchunk, int nkch,
                                                   int rhs_mode, const double* __restrict__ c2f,
                                                   double* __restrict__ divout, int div32,
                                                   const double* __restrict__ tl_dc2, const double* __restrict__ tl_div,
                                                   S* __restrict__ ecols_a, const EnsMember* __restrict__ M,
                                                   const int4* __restrict__ items, int nitems, int cur,
                                                   S* __restrict__ qwo_a, double beta_r) {
  constexpr bool NOQ = (XF & kStoreNoQ) != 0, RECON = (XF & kStageRecon) != 0;
  static_assert(XF == 0 || (MODE == 0 && sizeof(S) == 8), "the stage schedule variants are FP64, single grid");
  static_assert(!RECON || (KIND == 1 && !NOQ), "a rebuilt input is a middle stage's");
  static_assert(!NOQ || KIND != 2, "the last stage stores q");
  __shared__ double sp[2][TJ + 6][TI + 6];
  __shared__ double svx[2][TJ][TI + 6];
  __shared__ double sve[2][TJ + 6][TI];

  constexpr bool ENS = MODE == 1, TL = MODE == 2;
  const int n_items = ENS ? nitems : 1;
  for (int item = ENS ? (int)blockIdx.x : 0; item < n_items; item += ENS ? (int)gridDim.x : 1) {
  int bi, bj, P, k0, k1, mem = 0;
  if constexpr (ENS) {
    const int4 w = items[item];
    mem = w.x;
    bi = w.y & 0xffff;
    bj = w.y >> 16;
    P = w.z & 0xff;
    k0 = w.z >> 8;
    k1 = w.w;
  } else {
    const int nti = (Targ.n + TI - 1) / TI, ntj = (Targ.n + TJ - 1) / TJ;
    const int nb = nti * ntj * 6 * nkch;
    int b = xcd_remap(blockIdx.x, nb);
    bi = b % nti;
    b /= nti;
    bj = b % ntj;
    b /= ntj;
    P = b % 6;
    const int ch = b / 6;
    k0 = ch * kchunk;
    k1 = min(k0 + kchunk, Targ.nr);
    if (k0 >= Targ.nr) return;
  }
  const DevTables& T = ENS ? M[mem].T : Targ;
  const auto qin = to_global(ENS ? reinterpret_cast<const S*>(M[mem].q[cur]) : qin_a);
  const auto qout = to_global(ENS ? reinterpret_cast<S*>(M[mem].q[1 - cur]) : qout_a);
  const auto qw = to_global(ENS ? reinterpret_cast<S*>(M[mem].qw) : qw_a);
  const auto gxp = to_global(ENS ? M[mem].gxp : gxp_a);
  const auto gyp = to_global(ENS ? M[mem].gyp : gyp_a);
  const auto gxv = to_global(ENS ? M[mem].gxv : gxv_a);
  const auto gyv = to_global(ENS ? M[mem].gyv : gyv_a);
  const double dt = ENS ? M[mem].dt : dt_a;
  const int n = T.n, nr = T.nr;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = bi * TI, j0 = bj * TJ;
  const int i = i0 + tx, j = j0 + ty;
  // own point inside the panel (always, unless PART); outside threads load
  // through the nearest in-panel point and store nothing
  const bool own_in = !PART || (i < n && j < n);
  const int ic = PART ? min(i, n - 1) : i, jc = PART ? min(j, n - 1) : j;

  const size_t plane = (size_t)n * n, nf = plane * nr;
  const size_t ij = (size_t)jc * n + ic;
  const auto f0 = qin + (size_t)(P * 5) * nf;  // rho'
  const auto f1 = f0 + nf;                     // v_xi
  const auto f2 = f1 + nf;                     // v_eta
  const auto f3 = f2 + nf;                     // v_r
  const auto f4 = f3 + nf;                     // p'
  const auto o0 = qout + (size_t)(P * 5) * nf;
  const auto w0 = qw + (size_t)(P * 5) * nf;
  // where qw is stored: in place, or (RECON) in the other accumulator
  const auto wo0 = RECON ? to_global(qwo_a) + (size_t)(P * 5) * nf : w0;
  // edge-column copy: layer l = i (W) or n-1-i (E); the others store none
  const int ecl = i <= 3 ? i : n - 1 - i;
  S* const ecols = ENS ? reinterpret_cast<S*>(M[mem].ecols) : ecols_a;
  const bool ec_on = !rhs_mode && ecols != nullptr && ecl >= 1 && ecl <= 3 && own_in;
  const auto ecw = to_global(ecols);  // stored through only when ec_on
  const size_t ec_base = ec_on ? (size_t)(P * 2 + (i > 3)) * nr * edge_cols_plane(n) + (size_t)j * 3 + (ecl - 1) : 0;

  // Halo cells this thread fills: source pointer and plane stride, from the
  // field inside the panel or from the gathered ghost strips beyond its edge.
  const size_t gx_side = (size_t)nr * n * 3, gy_side = (size_t)nr * 3 * n;
  const auto gxpW = gxp + (size_t)(P * 2) * gx_side;
  const auto gxvW = gxv + (size_t)(P * 2) * gx_side;
  const auto gypS = gyp + (size_t)(P * 2) * gy_side;
  const auto gyvS = gyv + (size_t)(P * 2) * gy_side;
  const bool hx = tx < 3 || tx >= TI - 3, hy0 = ty < 3, hy1 = ty >= TJ - 3;
  // Each halo slot: a 32-bit offset inside one plane of its source and
  // whether the source is a ghost strip (plane stride 3n) or the field
  // (plane stride n*n); the p' and velocity halos share the offset.
  int hx_off = 0, hy0_off = 0, hy1_off = 0;
  bool hx_g = false, hy0_g = false, hy1_g = false;
  int hx_col = 0;
  // PART: a slot whose cell lies beyond the panel and beyond its three ghost
  // layers, or in a row or column outside the panel, keeps the fixed address
  // (no stencil of an in-panel point reads it).
  if (hx) {
    const bool lo = tx < 3;
    const int ii = lo ? i0 - 3 + tx : i + 3;
    hx_col = lo ? tx : tx + 6;
    if (PART && (j >= n || ii >= n + 3)) {
    } else if (ii >= 0 && ii < n) {
      hx_off = j * n + ii;
    } else {
      hx_g = true;
      hx_off = lo ? j * 3 + (-ii - 1) : (int)gx_side + j * 3 + (ii - n);
    }
  }
  if (hy0) {
    const int jj = j0 - 3 + ty;
    if (PART && i >= n) {
    } else if (jj >= 0) {
      hy0_off = jj * n + i;
    } else {
      hy0_g = true;
      hy0_off = (-jj - 1) * n + i;
    }
  }
  if (hy1) {
    const int jj = j + 3;
    if (PART && (i >= n || jj >= n + 3)) {
    } else if (jj < n) {
      hy1_off = jj * n + i;
    } else {
      hy1_g = true;
      hy1_off = (int)gy_side + (jj - n) * n + i;
    }
  }
  // PART: an outside thread within three points of the E edge (in a panel
  // row) or of the N edge (in a panel column) owns an LDS cell that holds
  // ghosts: E ghosts of p' and v_xi, or N ghosts of p' and v_eta.
  int og = 0, og_off = 0;
  auto ogp_src = gxpW, ogv_src = gxvW;  // without ghosts: a fixed in-bounds address
  if (PART) {
    if (j < n && i >= n && i < n + 3) {
      og = 1;
      ogp_src = gxpW + gx_side;
      ogv_src = gxvW + gx_side;
      og_off = j * 3 + (i - n);
    } else if (i < n && j >= n && j < n + 3) {
      og = 2;
      ogp_src = gypS + gy_side;
      ogv_src = gyvS + gy_side;
      og_off = (j - n) * n + i;
    }
  }

  constexpr bool kWide = sizeof(S) == sizeof(double);
  // The loads of a plane carry no branches (every halo slot reads a valid
  // address: offset 0 of the plane when the thread has no such halo cell), so
  // the waits before their use count exactly the stores issued after them.
  // FP32 state: field halo cells load as FP32, ghost-strip cells as FP64 in
  // waves that hold one (a wave-uniform branch).
  const bool wg_x = kWide || __ballot(hx_g) != 0, wg_y0 = kWide || __ballot(hy0_g) != 0,
             wg_y1 = kWide || __ballot(hy1_g) != 0;
  const auto c2base = to_global(c2f ? c2f : T.ci);  // without c2f: a fixed in-bounds address
  auto load_plane = [&](int k, StagePlane<S>& L) {
    const size_t kp = (size_t)k * plane + ij;
    L.vx = f1[kp];
    L.ve = f2[kp];
    L.q1 = __builtin_nontemporal_load(f0 + kp);
#pragma unroll
    for (int f = 0; f < 5; ++f)
      L.w[f] = (KIND == 0) ? S(0) : __builtin_nontemporal_load(w0 + f * nf + kp);
    const size_t kf = (size_t)k * plane, kg = (size_t)k * 3 * n;
    if constexpr (kWide) {
      L.hpx = hx_g ? gxpW[kg + hx_off] : f4[kf + hx_off];
      L.hvx = hx_g ? gxvW[kg + hx_off] : f1[kf + hx_off];
      L.hpy0 = hy0_g ? gypS[kg + hy0_off] : f4[kf + hy0_off];
      L.hve0 = hy0_g ? gyvS[kg + hy0_off] : f2[kf + hy0_off];
      L.hpy1 = hy1_g ? gypS[kg + hy1_off] : f4[kf + hy1_off];
      L.hve1 = hy1_g ? gyvS[kg + hy1_off] : f2[kf + hy1_off];
      if constexpr (RECON) {  // qw of the field cells (a ghost cell reads offset 0 and does not use it)
        const int fx = hx_g ? 0 : hx_off, fy0 = hy0_g ? 0 : hy0_off, fy1 = hy1_g ? 0 : hy1_off;
        L.rpx = w0[4 * nf + kf + fx];
        L.rvx = w0[nf + kf + fx];
        L.rpy0 = w0[4 * nf + kf + fy0];
        L.rve0 = w0[2 * nf + kf + fy0];
        L.rpy1 = w0[4 * nf + kf + fy1];
        L.rve1 = w0[2 * nf + kf + fy1];
      }
    } else {
      const int fx = hx_g ? 0 : hx_off, fy0 = hy0_g ? 0 : hy0_off, fy1 = hy1_g ? 0 : hy1_off;
      L.fpx = f4[kf + fx];
      L.fvx = f1[kf + fx];
      L.fpy0 = f4[kf + fy0];
      L.fve0 = f2[kf + fy0];
      L.fpy1 = f4[kf + fy1];
      L.fve1 = f2[kf + fy1];
      if (wg_x) {
        const int gx = hx_g ? hx_off : 0;
        L.hpx = gxpW[kg + gx];
        L.hvx = gxvW[kg + gx];
      }
      if (wg_y0) {
        const int gy = hy0_g ? hy0_off : 0;
        L.hpy0 = gypS[kg + gy];
        L.hve0 = gyvS[kg + gy];
      }
      if (wg_y1) {
        const int gy = hy1_g ? hy1_off : 0;
        L.hpy1 = gypS[kg + gy];
        L.hve1 = gyvS[kg + gy];
      }
    }
    if constexpr (PART) {
      L.ogp = ogp_src[kg + og_off];
      L.ogv = ogv_src[kg + og_off];
    }
    // plane entering the radial window when the window slides before k
    const int kn = min(k + 3, nr - 1);
    L.pn = f4[(size_t)kn * plane + ij];
    L.rn = f3[(size_t)kn * plane + ij];
    if constexpr (RECON) {
      L.wpn = w0[4 * nf + (size_t)kn * plane + ij];
      L.wrn = w0[3 * nf + (size_t)kn * plane + ij];
    }
    // per-point inputs of the RHS, prefetched with the plane: a load issued
    // in the body would be waited on together with the whole prefetch
    const size_t dx = (size_t)P * nf + kp;
    L.c2 = c2base[c2f ? dx : 0];
    if constexpr (TL) {
      L.tdc2 = tl_dc2[dx];
      L.tdiv = tl_div[dx];
    }
  };

  const double Ci = to_global(T.ci)[ic], Dj = to_global(T.dj)[jc];
  const double yCi = 1.0 / Ci, yDj = 1.0 / Dj;  // the divisors fixed for this thread's column
  // Reciprocal-based quotients in the FP32-state stages only (c3 stage 2.87 ->
  // 2.74 ms with a guard per quotient, and a further 1% with one guard per
  // group); the FP64 stages divide in hardware (div_group).
  const bool fdiv = T.fast_div != 0;
  const double dc = to_global(T.dc)[ij], tdx = to_global(T.tdx)[ij], tde = to_global(T.tde)[ij],
               xy = to_global(T.xy)[ij];

  // Radial windows of p' and v_r over planes kb..kb+6 (storage type).
  S pw[7], rw[7];
  int kb = clampi(k0 - 3, 0, nr - 7);
#pragma unroll
  for (int q = 0; q < 7; ++q) {  // (nr == 1, the 2-D case: every slot holds plane 0, with DDR weights 0)
    const int kq = min(kb + q, nr - 1);
    pw[q] = f4[(size_t)kq * plane + ij];
    rw[q] = f3[(size_t)kq * plane + ij];
    if constexpr (RECON) {
      pw[q] = pw[q] + beta_r * w0[4 * nf + (size_t)kq * plane + ij];
      rw[q] = rw[q] + beta_r * w0[3 * nf + (size_t)kq * plane + ij];
    }
  }
  // The next plane is loaded while this one is computed (a second plane in
  // flight was measured: no gain for FP64 or FP32 storage).  Each iteration
  // ends by moving the prefetched registers into place, which waits for
  // every vector-memory operation issued before it (the counter retires in
  // issue order).  So a plane's results are held in registers and stored at
  // the top of the next iteration, before its prefetch: by the end of that
  // iteration the stores have had a whole plane's computation to retire.
  StagePlane<S> nxt;
  load_plane(k0, nxt);
  S pq0, pq1, pq2, pq3, pq4, pw0, pw1, pw2, pw3, pw4;  // the previous plane's q, qw (or RHS)
  size_t px = 0;
  int pk = 0;  // the plane of px
  auto flush = [&]() {
    const S pq[5] = {pq0, pq1, pq2, pq3, pq4}, pwv[5] = {pw0, pw1, pw2, pw3, pw4};
    if (!own_in) {
    } else if (rhs_mode) {
#pragma unroll
      for (int f = 0; f < 5; ++f) o0[f * nf + px] = pq[f];
    } else {
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        if (!NOQ) __builtin_nontemporal_store(pq[f], o0 + f * nf + px);
        if (KIND != 2) __builtin_nontemporal_store(pwv[f], wo0 + f * nf + px);
      }
      if (ec_on) {
        const size_t x = ec_base + (size_t)pk * edge_cols_plane(n);
#pragma unroll
        for (int f = 0; f < 5; ++f)
          if (KIND == 2 || f == 1 || f == 2 || f == 4) ecw[x + (size_t)f * 3 * n] = pq[f];
      }
    }
  };

  for (int k = k0; k < k1; ++k) {
    const StagePlane<S> cur = nxt;
    if (k > k0) flush();
    load_plane(min(k + 1, k1 - 1), nxt);  // the last plane's prefetch is loaded and not used
    const int want = clampi(k - 3, 0, nr - 7);
    if (want != kb) {  // slides by exactly one plane: plane k+3 enters
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        pw[q] = pw[q + 1];
        rw[q] = rw[q + 1];
      }
      pw[6] = RECON ? cur.pn + beta_r * cur.wpn : cur.pn;
      rw[6] = RECON ? cur.rn + beta_r * cur.wrn : cur.rn;
      kb = want;
    }
    // p', v_r at plane k: window slot k - kb (3 on interior planes)
    const int wk = k - kb;
    S pcs = pw[0], rcs = rw[0];
#pragma unroll
    for (int q = 1; q < 7; ++q) {
      pcs = wk == q ? pw[q] : pcs;
      rcs = wk == q ? rw[q] : rcs;
    }
    const double pc = pcs, rc = rcs;
    const int bf = k & 1;
    // the stage input at the own point (RECON: rebuilt from q and qw)
    const double vx = RECON ? cur.vx + beta_r * cur.w[1] : cur.vx;
    const double ve = RECON ? cur.ve + beta_r * cur.w[2] : cur.ve;
    const double q1 = RECON ? cur.q1 + beta_r * cur.w[0] : cur.q1;
    double own_p = pc, own_vx = vx, own_ve = ve;
    if constexpr (PART) {
      if (og == 1) {
        own_p = cur.ogp;
        own_vx = cur.ogv;
      } else if (og == 2) {
        own_p = cur.ogp;
        own_ve = cur.ogv;
      }
    }
    sp[bf][ty + 3][tx + 3] = own_p;
    svx[bf][ty][tx + 3] = own_vx;
    sve[bf][ty + 3][tx] = own_ve;
    if constexpr (RECON) {  // field halo cells rebuilt; ghost cells are the gathered stage input
      if (hx) {
        sp[bf][ty + 3][hx_col] = hx_g ? cur.hpx : cur.hpx + beta_r * cur.rpx;
        svx[bf][ty][hx_col] = hx_g ? cur.hvx : cur.hvx + beta_r * cur.rvx;
      }
      if (hy0) {
        sp[bf][ty][tx + 3] = hy0_g ? cur.hpy0 : cur.hpy0 + beta_r * cur.rpy0;
        sve[bf][ty][tx] = hy0_g ? cur.hve0 : cur.hve0 + beta_r * cur.rve0;
      }
      if (hy1) {
        sp[bf][ty + 6][tx + 3] = hy1_g ? cur.hpy1 : cur.hpy1 + beta_r * cur.rpy1;
        sve[bf][ty + 6][tx] = hy1_g ? cur.hve1 : cur.hve1 + beta_r * cur.rve1;
      }
    } else {
      if (hx) {
        sp[bf][ty + 3][hx_col] = kWide || hx_g ? cur.hpx : static_cast<double>(cur.fpx);
        svx[bf][ty][hx_col] = kWide || hx_g ? cur.hvx : static_cast<double>(cur.fvx);
      }
      if (hy0) {
        sp[bf][ty][tx + 3] = kWide || hy0_g ? cur.hpy0 : static_cast<double>(cur.fpy0);
        sve[bf][ty][tx] = kWide || hy0_g ? cur.hve0 : static_cast<double>(cur.fve0);
      }
      if (hy1) {
        sp[bf][ty + 6][tx + 3] = kWide || hy1_g ? cur.hpy1 : static_cast<double>(cur.fpy1);
        sve[bf][ty + 6][tx] = kWide || hy1_g ? cur.hve1 : static_cast<double>(cur.fve1);
      }
    }
    __syncthreads();

    // 7-point DRP stencils, ascending taps from 0 (derivative.f90:137-148,
    // 456-467, 656-668, 704-754).  In xi and eta the coefficients are the
    // same (d_eta = d_xi) and the centre tap is 0 (both checked when the
    // tables are uploaded): its product, +-0 for a finite value, leaves the
    // sum unchanged, so it is not formed.  Fewer uniform operands also keep
    // the kernel's scalar registers from spilling.
    double dvx = 0.0, dve = 0.0, dpx = 0.0, dpe = 0.0, dvr = 0.0, dpr = 0.0;
    // the radial row is block-uniform and read-only: through the constant
    // address space it comes in with scalar loads, which the vector-load
    // waits do not count
    const auto& row = const_rows(T.rad)[k];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      if (q != 3) {
        dvx = dvx + T.fcx[q] * svx[bf][ty][tx + q];
        dve = dve + T.fcx[q] * sve[bf][ty + q][tx];
        dpx = dpx + T.fcx[q] * sp[bf][ty + 3][tx + q];
        dpe = dpe + T.fcx[q] * sp[bf][ty + q][tx + 3];
      }
      dvr = dvr + row.w[q] * rw[q];
      dpr = dpr + row.w[q] * pw[q];
    }

    // grad.f90:125-133 and :49-57.
    // quotients by C, D, r and rho0 (grad.f90:52-53, 128-130;
    // compute_rhs_acoustic.f90:88-146), in two groups: the second divides
    // the first's results
    double qa[7];
    {
      const double na[7] = {dvx, dve, 2.0 * rc, xy * dpe, xy * dpx, dpr, q1 * row.g0};
      const double ba[7] = {Ci, Dj, row.r, Dj, Ci, row.rho0, row.rho0};
      const double ya[7] = {yCi, yDj, row.inv_r, yDj, yCi, row.inv_rho0, row.inv_rho0};
      div_group<kWide ? false : true>(na, ba, ya, fdiv, qa);
    }
    const double div = (row.inv_r * ((dc * (qa[0] + qa[1]) - tdx * vx) - tde * ve) + dvr) + qa[2];
    const double gxi = row.inv_r * (Dj * dpx + qa[3]);
    const double get = row.inv_r * (Ci * dpe + qa[4]);

    // compute_rhs_acoustic.f90:38-149 (forcing omitted).
    double r1 = -(row.rho0 * div);
    r1 = r1 - rc * row.drho0dr;
    double qb[2];
    {
      const double nb[2] = {gxi, get};
      const double bb[2] = {row.rho0, row.rho0};
      const double yb[2] = {row.inv_rho0, row.inv_rho0};
      div_group<kWide ? false : true>(nb, bb, yb, fdiv, qb);
    }
    double r2 = -qb[0];
    double r3 = -qb[1];
    double r4 = -qa[5];
    r4 = r4 - qa[6];
    // c2f: optional 3-D c^2 (n,n,nr,6); the radial profile otherwise
    const double c2rho0 = c2f ? cur.c2 * row.rho0 : row.c2rho0;
    const size_t dx = (size_t)P * nf + (size_t)k * plane + ij;
    if (divout && own_in) {  // div(v) of the stage input, kept for the adjoint (FP64, or FP32 when div32)
      if (div32)
        reinterpret_cast<float*>(divout)[dx] = static_cast<float>(div);
      else
        divout[dx] = div;
    }
    double r5 = -(c2rho0 * div);
    // tangent-linear run: d r5 / d c^2 = -rho0 div of the primal stage input
    if constexpr (TL) r5 = r5 - (cur.tdc2 * row.rho0) * cur.tdiv;
    r5 = r5 + row.rho0g0 * rc;

    // no BC_RHS and no SPONGE in 2-D (bc_acoustic_wave.f90:11, sponge.f90:18)
    const bool do_bc = (rhs_mode ? (rhs_mode & 1) : true) && nr > 1;
    const bool do_sp = (rhs_mode ? (rhs_mode & 2) : true) && nr > 1;
    if (do_bc && (k == 0 || k == nr - 1)) {  // bc_acoustic_wave.f90:14-34
      r1 = 0.0; r2 = 0.0; r3 = 0.0; r4 = 0.0; r5 = 0.0;
    }
    if (do_sp) {  // sponge.f90:48-50, after BC_RHS (timestep.f90:65-66)
      const double sg = row.sigma;
      r1 = r1 - sg * (q1 - 0.0);
      r4 = r4 - sg * (rc - 0.0);
      r5 = r5 - sg * (pc - 0.0);
    }

    px = (size_t)k * plane + ij;
    pk = k;
    if (rhs_mode) {
      pq0 = static_cast<S>(r1);
      pq1 = static_cast<S>(r2);
      pq2 = static_cast<S>(r3);
      pq3 = static_cast<S>(r4);
      pq4 = static_cast<S>(r5);
    } else {
      // timestep.f90:68-69: qw = alpha*qw + dt*rhs ; q = q + beta*qw
      const double w0n = alpha * cur.w[0] + dt * r1, w1n = alpha * cur.w[1] + dt * r2,
                   w2n = alpha * cur.w[2] + dt * r3, w3n = alpha * cur.w[3] + dt * r4,
                   w4n = alpha * cur.w[4] + dt * r5;
      pq0 = static_cast<S>(q1 + beta * w0n);
      pq1 = static_cast<S>(vx + beta * w1n);
      pq2 = static_cast<S>(ve + beta * w2n);
      pq3 = static_cast<S>(rc + beta * w3n);
      pq4 = static_cast<S>(pc + beta * w4n);
      pw0 = static_cast<S>(w0n);
      pw1 = static_cast<S>(w1n);
      pw2 = static_cast<S>(w2n);
      pw3 = static_cast<S>(w3n);
      pw4 = static_cast<S>(w4n);
    }
  }
  flush();
  if constexpr (ENS) __syncthreads();  // the next item rewrites the LDS tiles
  }
}

// FILTER_EXP for one field of one panel over a chunk of radial planes.
//   fgx/fgy: ghosts of the unfiltered field (filter.f90:25-28), layouts as in
//   k_gather, 12 sides (6 panels x 2) per field.
// Per plane: the (TJ+6) x (TI+6) tile of the unfiltered field goes to LDS
// (prefetched into registers two planes ahead), FILTER_XI runs on the
// in-panel rows, FILTER_ETA on the result, and FILTER_R accumulates the
// plane's terms into the eleven pending outputs in registers.
template <int TI, int TJ>
struct FilterLds {
  alignas(16) double sq[TJ + 6][TI + 6];  // rows of TI + 6 doubles: 16-byte aligned pairs
  alignas(16) double sx[TJ + 6][TI];
};

// The tile (bi, bj) of field F of panel P (PF = 5P + F) over planes [k0, k1).
// PART: the tile may reach past the panel edge; threads past it store nothing,
// and tile cells beyond the panel's three ghost layers stay 0.
// PAIR: FILTER_XI tasks of two adjacent outputs (below).
template <int TI, int TJ, typename S, bool PART = false, bool PAIR = (sizeof(S) == 4)>
__device__ __forceinline__ void filter_tile(const DevTables& T, int bi, int bj, int PF, int k0, int k1,
                                            const S* __restrict__ qin, S* __restrict__ qout,
                                            const double* __restrict__ fgx, const double* __restrict__ fgy,
                                            FilterLds<TI, TJ>& lds) {
  constexpr int NT = TI * TJ;
  constexpr int NCELL = (TJ + 6) * (TI + 6);
  constexpr int NC = (NCELL + NT - 1) / NT;
  auto& sq = lds.sq;
  auto& sx = lds.sx;
  const int n = T.n, nr = T.nr;
  const int P = PF / 5, F = PF % 5;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = bi * TI, j0 = bj * TJ;
  const int i = i0 + tx, j = j0 + ty;
  const bool own_in = !PART || (i < n && j < n);

  const size_t plane = (size_t)n * n, nf = plane * nr;
  const S* g = qin + (size_t)PF * nf;
  S* o = qout + (size_t)PF * nf;
  const size_t gx_side = (size_t)nr * n * 3, gy_side = (size_t)nr * 3 * n;
  const size_t fstride = (size_t)12 * gx_side;
  const double* gxW = fgx + F * fstride + (size_t)(P * 2) * gx_side;
  const double* gyS = fgy + F * fstride + (size_t)(P * 2) * gy_side;
  const size_t ij = own_in ? (size_t)j * n + i : 0;

  // 2-D case (nr == 1): FILTER_R returns its input (filter.f90:647-650)
  const bool twod = nr == 1;
  // filter.f90:652,726-777: the five planes at each end are zero.
  if (own_in && !twod)
    for (int k = k0; k < k1; ++k)
      if (k < kNstr || k >= nr - kNstr) o[(size_t)k * plane + ij] = static_cast<S>(0.0);
  const int a = twod ? 0 : max(k0, kNstr), bnd = twod ? 1 : min(k1, nr - kNstr);
  if (a >= bnd) return;

  // Tile cells this thread fills: kind 1 a field cell, 2 a ghost-strip cell,
  // 0 none.  Field cells load from the plane base (a wave-uniform address)
  // plus a per-lane 32-bit byte offset, so advancing a plane costs no vector
  // instruction; ghost cells (plane stride 3n, uniform too) have per-lane
  // pointers.  Every load is issued unconditionally from a valid address:
  // lanes without a field cell read the plane's first element, lanes without
  // a ghost cell the strip's first element.  Kind-0 cells (the corners beyond
  // the panel in both directions, cells past a partial tile's ghost layers)
  // feed only outputs that are not stored, so they keep whatever was loaded.
  const int tid = ty * TI + tx;
  unsigned foff[NC];
  const double* gp[NC];
  int kind[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int e = tid + c * NT;
    kind[c] = 0;
    foff[c] = 0;
    gp[c] = gxW;
    if (e < NCELL) {
      const int rr = e / (TI + 6), cc = e % (TI + 6);
      const int gj = j0 - 3 + rr, gi = i0 - 3 + cc;
      const bool jin = gj >= 0 && gj < n, iin = gi >= 0 && gi < n;
      if (jin && iin) {
        foff[c] = (unsigned)(((size_t)gj * n + gi) * sizeof(S));  // < n*n*sizeof(S)
        kind[c] = 1;
      } else if (jin && gi < n + 3) {  // gi >= -3 always; beyond the E ghosts: none
        gp[c] = (gi < 0) ? gxW + (size_t)gj * 3 + (-gi - 1) : gxW + gx_side + (size_t)gj * 3 + (gi - n);
        kind[c] = 2;
      } else if (iin && gj < n + 3) {
        gp[c] = (gj < 0) ? gyS + (size_t)(-gj - 1) * n + gi : gyS + gy_side + (size_t)(gj - n) * n + gi;
        kind[c] = 2;
      }
    }
  }
  // Two planes in flight in alternating buffers pa / pb: a plane is consumed
  // whole by the LDS write that starts its step, and its buffer is refilled
  // with the plane two ahead right there, so a load has the rest of its step
  // and the whole next step to arrive (no register copies between
  // iterations, which would wait on the loads).  Field cells stay in the
  // storage type in flight (.f) and are widened when they go to LDS, next to
  // the FP64 ghost-strip cells (.g): a conversion at the load would also wait
  // on every load in flight.
  // FP64 state: one value per cell, from the field or (in a wave with ghost
  // cells, by a per-lane address select) from the strip.
  constexpr bool kWide = sizeof(S) == sizeof(double);
  struct FPlaneN {
    double g[NC];
    S f[NC];
  };
  struct FPlaneW {
    double g[NC];
  };
  using FPlane = std::conditional_t<kWide, FPlaneW, FPlaneN>;
  // the ghost-strip load and the select only in waves that hold a ghost cell
  // (a wave-uniform branch; tiles off the panel edges have none)
  bool wave_g[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) wave_g[c] = __ballot(kind[c] == 2) != 0;
  const size_t gstride = (size_t)3 * n;
  // the loads name global memory explicitly (the pointers come from kernel
  // arguments through a select the compiler cannot see through)
  using GD = const __attribute__((address_space(1))) double*;
  using GS = const __attribute__((address_space(1))) S*;
  auto load_plane = [&](int kk, FPlane& dst) {
    const char* fb = reinterpret_cast<const char*>(g + (size_t)kk * plane);  // wave-uniform
    const size_t goff = (size_t)kk * gstride;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if constexpr (kWide) {
        if (wave_g[c]) {
          const double* p = kind[c] == 2 ? gp[c] + goff : reinterpret_cast<const double*>(fb + foff[c]);
          dst.g[c] = *(GD)p;
        } else {
          dst.g[c] = *(GD)(fb + foff[c]);
        }
      } else {
        dst.f[c] = *(GS)(fb + foff[c]);
        dst.g[c] = wave_g[c] ? *(GD)(gp[c] + goff) : 0.0;
      }
    }
  };
  auto widen = [&](const FPlane& L, int c) -> double {
    if constexpr (kWide) {
      return L.g[c];
    } else {
      const double fv = static_cast<double>(L.f[c]);
      if (wave_g[c]) return kind[c] == 2 ? L.g[c] : fv;
      return fv;
    }
  };

  // FILTER_XI on in-panel rows (filter.f90:176-315); rows beyond the panel
  // edge keep the unfiltered ghost, which FILTER_ETA uses (filter.f90:43).
  // The sums start from the first product, not from the reference's 0.0: the
  // two differ only in the sign of a zero result.  A sum's value does not
  // depend on the signs of zero terms, so FILTER_ETA's outputs keep the
  // reference's values (possibly with another zero sign), and the stored
  // sum -- FILTER_R's, or FILTER_ETA's in 2-D, both from 0.0, which never
  // gives -0 -- has the reference's bits.
  // PAIR: a task is two adjacent outputs of one row, read as four aligned
  // 16-byte pairs (eight values for two 7-tap sums), 22 x 16 tasks for a
  // 32 x 16 tile: half the LDS reads of one output per thread (FP32 filter
  // at c3 2.79 -> 2.61 ms).  The FP64 filter takes pairs on panels up to
  // 128 (filter 6x128^2x256 -3%, 6x96^2x192 -7%, the ensemble step -0.8%)
  // and one output per thread on larger ones (6x256^2x384: 2.93 -> 3.14 ms
  // with pairs).
  typedef double dpair __attribute__((ext_vector_type(2)));  // one 16-byte LDS access
  auto xi_pass = [&]() {
    if constexpr (!PAIR) {
      for (int rr = ty; rr < TJ + 6; rr += TJ) {
        const int gj = j0 - 3 + rr;
        double xv;
        if (gj >= 0 && gj < n) {
          xv = T.filt[0] * sq[rr][tx];
#pragma unroll
          for (int q = 1; q < 7; ++q) xv = xv + T.filt[q] * sq[rr][tx + q];
        } else {
          xv = sq[rr][tx + 3];
        }
        sx[rr][tx] = xv;
      }
      return;
    }
    constexpr int NTASK = (TJ + 6) * (TI / 2);
    for (int t = tid; t < NTASK; t += NT) {
      const int rr = t / (TI / 2), c0 = 2 * (t % (TI / 2));
      const dpair* src = reinterpret_cast<const dpair*>(&sq[rr][c0]);
      double v[8];
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const dpair d = src[h];
        v[2 * h] = d.x;
        v[2 * h + 1] = d.y;
      }
      const int gj = j0 - 3 + rr;
      double x0, x1;
      if (gj >= 0 && gj < n) {
        x0 = T.filt[0] * v[0];
        x1 = T.filt[0] * v[1];
#pragma unroll
        for (int q = 1; q < 7; ++q) {
          x0 = x0 + T.filt[q] * v[q];
          x1 = x1 + T.filt[q] * v[q + 1];
        }
      } else {
        x0 = v[3];
        x1 = v[4];
      }
      dpair out;
      out.x = x0;
      out.y = x1;
      *reinterpret_cast<dpair*>(&sx[rr][c0]) = out;
    }
  };

  if (twod) {  // the one plane: FILTER_XI, FILTER_ETA, stored as they come
    FPlane p0;
    load_plane(0, p0);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int e = tid + c * NT;
      if (e < NCELL) (&sq[0][0])[e] = widen(p0, c);
    }
    __syncthreads();
    xi_pass();
    __syncthreads();
    double v = 0.0;  // stored: the reference's sum from 0.0
#pragma unroll
    for (int q = 0; q < 7; ++q) v = v + T.filt[q] * sx[ty + q][tx];
    if (own_in) o[ij] = static_cast<S>(v);
    return;
  }

  // FILTER_R (filter.f90:682-693), output plane k = sum over q of
  // filtr[q] * eta(k - 5 + q), summed from 0.0 in q order.  The taps are
  // symmetric (filtr[q] == filtr[10 - q], checked when the tables are
  // uploaded), so each arriving plane's six distinct products feed the eleven
  // pending outputs: acc[d] is output plane kk - 5 + d, which takes the term
  // q = 10 - d of plane kk, the same order and the same products as the
  // sum over a window of the last eleven planes with 6 multiplies instead of 11.
  double acc[11];
#pragma unroll
  for (int q = 0; q < 11; ++q) acc[q] = 0.0;

  const int kbeg = a - kNstr, kend = bnd + kNstr;  // kend - kbeg >= 11
  // store: the step completes an output plane (every step from the 11th on)
  auto step = [&](int kk, FPlane& cur, auto store) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int e = tid + c * NT;
      if (e < NCELL) (&sq[0][0])[e] = widen(cur, c);
    }
    load_plane(min(kk + 2, kend - 1), cur);  // past the end: a plane that is loaded and not used
    __syncthreads();
    xi_pass();
    __syncthreads();
    // FILTER_ETA (filter.f90:448-629), from the first product, as FILTER_XI:
    // FILTER_R's sum from 0.0 does not see the sign of a zero input.
    double v = T.filt[0] * sx[ty][tx];
#pragma unroll
    for (int q = 1; q < 7; ++q) v = v + T.filt[q] * sx[ty + q][tx];
    double pr[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) pr[m] = T.filtr[m] * v;
#pragma unroll
    for (int d = 0; d < 10; ++d) acc[d] = acc[d] + pr[d < 6 ? d : 10 - d];
    acc[10] = 0.0 + pr[0];
    if constexpr (decltype(store)::value) {
      if (own_in) o[(size_t)(kk - kNstr) * plane + ij] = static_cast<S>(acc[0]);
    }
#pragma unroll
    for (int d = 0; d < 10; ++d) acc[d] = acc[d + 1];
  };
  // Branch-free memory traffic per iteration (a fixed warm-up, then pairs of
  // storing steps, then at most one more), so the waits before each LDS
  // write leave the other buffer's loads in flight.
  using NoStore = std::integral_constant<bool, false>;
  using Store = std::integral_constant<bool, true>;
  FPlane pa, pb;
  load_plane(kbeg, pa);
  load_plane(kbeg + 1, pb);
  int kk = kbeg;
  for (; kk < kbeg + 2 * kNstr; kk += 2) {  // fill the radial window: planes a-5 .. a+4
    step(kk, pa, NoStore());
    step(kk + 1, pb, NoStore());
  }
  for (; kk + 1 < kend; kk += 2) {
    step(kk, pa, Store());
    step(kk + 1, pb, Store());
  }
  if (kk < kend) step(kk, pa, Store());
}

// Occupancy floor: six waves per SIMD (80 registers), three 512-thread
// blocks per CU.  Left alone, the FP32-state filter takes 84 registers and
// runs five waves, two blocks per CU, which leaves the barriers of each
// plane less to overlap with; at 80 it does not spill (the 32x16 and 16x16
// tiles; the other shapes spill 36 bytes and keep no floor).
template <int TI, int TJ, typename S>
constexpr int filter_waves() { return (sizeof(S) == 4 && TI * TJ >= 256 && !(TI == 64 && TJ == 4) && !(TI == 32 && TJ == 8)) ? 6 : 1; }

template <int TI, int TJ, typename S, bool PART = false, bool PAIR = (sizeof(S) == 4)>
__global__ __launch_bounds__(TI* TJ) __attribute__((amdgpu_waves_per_eu(filter_waves<TI, TJ, S>())))
void k_filter(DevTables T, const S* __restrict__ qin,
                                                    S* __restrict__ qout, const double* __restrict__ fgx,
                                                    const double* __restrict__ fgy, int kchunk, int nkch) {
  __shared__ FilterLds<TI, TJ> lds;
  const int n = T.n, nr = T.nr;
  const int nti = (n + TI - 1) / TI, ntj = (n + TJ - 1) / TJ;
  const int nb = nti * ntj * 30 * nkch;
  int b = xcd_remap(blockIdx.x, nb);
  const int bi = b % nti;
  b /= nti;
  const int bj = b % ntj;
  b /= ntj;
  const int PF = b % 30;  // panel*5 + field
  const int ch = b / 30;
  const int k0 = ch * kchunk;
  const int k1 = min(k0 + kchunk, nr);
  if (k0 >= nr) return;
  filter_tile<TI, TJ, S, PART, PAIR>(T, bi, bj, PF, k0, k1, qin, qout, fgx, fgy, lds);
}

// Ensemble FILTER_EXP: work items (member, tile, 5P+F, plane range), q[cur] -> q[1-cur].
// FILTER_XI in paired tasks (the members are small panels).  Six waves per
// SIMD (80 registers, a 32-byte spill at 32x16): faster than five waves
// without the spill (37.22 against 37.64 ms per 48-member step).
template <int TI, int TJ>
__global__ __launch_bounds__(TI* TJ) __attribute__((amdgpu_waves_per_eu(6))) void k_filter_ens(const EnsMember* __restrict__ M, const int4* __restrict__ items,
                                                        int nitems, int cur) {
  __shared__ FilterLds<TI, TJ> lds;
  for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
    const int4 w = items[item];
    const EnsMember& E = M[w.x];
    filter_tile<TI, TJ, double, false, true>(E.T, w.y & 0xffff, w.y >> 16, w.z & 0xff, w.z >> 8, w.w, E.q[cur], E.q[1 - cur],
                                E.fgx, E.fgy, lds);
    __syncthreads();  // the next item rewrites the LDS tiles
  }
}

// Ensemble gather over a group of members: their blocks laid end to end
// (table[x] = first block of the group's member x, 24 *
// gather_blocks_per_edge(nr, chunks) blocks each; table[nmem + x] = its index
// in M); blockIdx.y selects the field.
template <bool VEC, int NI>
__global__ __launch_bounds__(gather_threads_max(NI)) void k_gather_ens(const EnsMember* __restrict__ M, const int* __restrict__ table,
                                                     int nmem, int cur, int cols, int chunks) {
  extern __shared__ double sline[];
  const int b = blockIdx.x;
  int lo = 0, hi = nmem - 1;  // last member whose first block is <= b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (table[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  const EnsMember& E = M[table[nmem + lo]];
  const int li = b - table[lo], nr = E.T.nr, kc = gather_kc(nr, chunks);
  const int k0 = (li / 24) * kc, edge = li % 24, k1 = min(k0 + kc, nr);
  const long long fs = 12LL * nr * E.T.n * 3;
  const double* ec = cols ? E.ecols : nullptr;
  if (VEC)
    gather_edge<true, double, NI>(E.T, edge / 4, edge % 4, k0, k1, blockIdx.y, E.q[cur], ec, 4, E.gxp, E.gyp,
                                  E.gxv, E.gyv, fs, sline, nullptr, 0.0);
  else
    gather_edge<false, double, NI>(E.T, edge / 4, edge % 4, k0, k1, blockIdx.y, E.q[cur], ec, 0, E.fgx, E.fgy,
                                   nullptr, nullptr, fs, sline, nullptr, 0.0);
}

// Non-finite scan and v_r sum of squares; acc[0] counts non-finite values,
// acc[1] accumulates squares.
template <typename S>
__global__ void k_scan(const S* __restrict__ q, size_t total, size_t nf, double* acc, int mode, int field) {
  double s = 0.0;
  for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (size_t)gridDim.x * blockDim.x) {
    if (mode == 0) {
      const double v = static_cast<double>(q[x]);
      s += (isfinite(v) ? 0.0 : 1.0);
    } else {
      // `field` of panel p
      const size_t p = x / nf, r = x % nf;
      const double v = static_cast<double>(q[(p * 5 + field) * nf + r]);
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
template <typename S>
__global__ void k_place(const double* __restrict__ src, S* __restrict__ dst, size_t nf, int dfield) {
  const size_t total = nf * 6;
  for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (size_t)gridDim.x * blockDim.x) {
    const size_t p = x / nf, r = x % nf;
    dst[(p * 5 + dfield) * nf + r] = static_cast<S>(src[x]);
  }
}

// ---------------------------------------------------------------- launchers

static int pick_ti(int n) { return (n % 32 == 0) ? 32 : 16; }

int chunks_for(int tiles, int nr, int target_blocks, int min_chunk) {
  int nk = (target_blocks + tiles - 1) / tiles;
  const int maxk = nr / min_chunk > 0 ? nr / min_chunk : 1;
  if (nk > maxk) nk = maxk;
  if (nk < 1) nk = 1;
  return nk;
}

// Dynamic LDS of one gather launch; above 64 KiB the kernel is allowed it once.
template <typename K>
static hipError_t gather_lds(K kernel, size_t lds) {
  if (lds <= 64 * 1024) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(lds));
}

template <bool VEC, typename S, int NI>
static hipError_t launch_gather_v(const DevTables& T, const S* q, const S* ecols, int sfield, int nfields,
                                  double* gxs, double* gys, double* gxv, double* gyv, hipStream_t st, const S* qwr,
                                  double betar) {
  const long long field_stride = 12LL * T.nr * T.n * 3;
  const int kc = gather_kc(T.nr);
  dim3 grid((unsigned)(24 * ((T.nr + kc - 1) / kc)), (unsigned)nfields);
  const int nt = gather_threads(T.n, NI);
  const size_t lds = sizeof(double) * 2 * 3 * (VEC ? 3 : 1) * (size_t)T.n;  // n <= 512: at most 72 KiB
  hipError_t e = gather_lds(&k_gather<VEC, S, NI>, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_gather<VEC, S, NI>), grid, dim3(nt), lds, st, T, q, ecols, sfield, gxs, gys, gxv, gyv,
                     field_stride, qwr, betar);
  return hipGetLastError();
}

template <typename S>
static hipError_t launch_gather_t(const DevTables& T, const S* q, const S* ecols, bool vec, int sfield, int nfields,
                                  double* gxs, double* gys, double* gxv, double* gyv, hipStream_t st, const S* qwr,
                                  double betar) {
  const bool two = 3 * T.n > 1024;  // n > 341: two strip points per thread
  if (vec)
    return two ? launch_gather_v<true, S, 2>(T, q, ecols, sfield, nfields, gxs, gys, gxv, gyv, st, qwr, betar)
               : launch_gather_v<true, S, 1>(T, q, ecols, sfield, nfields, gxs, gys, gxv, gyv, st, qwr, betar);
  return two ? launch_gather_v<false, S, 2>(T, q, ecols, sfield, nfields, gxs, gys, gxv, gyv, st, qwr, betar)
             : launch_gather_v<false, S, 1>(T, q, ecols, sfield, nfields, gxs, gys, gxv, gyv, st, qwr, betar);
}

hipError_t launch_gather(const DevTables& T, const void* q, bool f32, bool vec, int sfield, int nfields, double* gxs,
                         double* gys, double* gxv, double* gyv, hipStream_t st, const void* ecols, const void* qwr,
                         double betar) {
  if (f32) {
    if (qwr) return hipErrorInvalidValue;  // rebuilt stage inputs: FP64 storage only
    return launch_gather_t(T, static_cast<const float*>(q), static_cast<const float*>(ecols), vec, sfield, nfields,
                           gxs, gys, gxv, gyv, st, static_cast<const float*>(nullptr), 0.0);
  }
  return launch_gather_t(T, static_cast<const double*>(q), static_cast<const double*>(ecols), vec, sfield, nfields,
                         gxs, gys, gxv, gyv, st, static_cast<const double*>(qwr), betar);
}

template <int TI, int TJ, typename S, bool PART = false>
static hipError_t launch_stage_t(const DevTables& T, const S* qin, S* qout, S* qw, const double* gxp,
                                 const double* gyp, const double* gxv, const double* gyv, int stage, double dt,
                                 int rhs_mode, hipStream_t st, const LaunchShape& ls, const double* c2f,
                                 double* divout, int div32, const double* tl_dc2, const double* tl_div, S* ecols,
                                 int xf, S* qwo, double beta_r) {
  const int tiles = ((T.n + TI - 1) / TI) * ((T.n + TJ - 1) / TJ) * 6;
  // radial chunks of >= 24 planes, ~12k blocks, at most 8 chunks: 8 chunks measured best at
  // 6x96^2x192 (0.342 ms per stage against 0.375 with 6), 6x128^2x256 and 6x256^2x384
  const int nkch = ls.stage_chunks > 0 ? ls.stage_chunks : std::min(8, chunks_for(tiles, T.nr, 12288, 24));
  const int kchunk = (T.nr + nkch - 1) / nkch;
  dim3 grid(tiles * nkch), block(TI, TJ);
  const double al = kAlphaRk4[stage < 0 ? 0 : stage], be = kBetaRk4[stage < 0 ? 0 : stage];
  const int kind = (rhs_mode || stage == 0) ? 0 : (stage < 4 ? 1 : 2);
#define SOLAR_STAGE_LAUNCH(K, M, X)                                                                                  \
  hipLaunchKernelGGL((k_stage<TI, TJ, K, S, M, PART, X>), grid, block, 0, st, T, qin, qout, qw, gxp, gyp, gxv, gyv,  \
                     al, be, dt, kchunk, nkch, rhs_mode, c2f, divout, div32, tl_dc2, tl_div, ecols, nullptr, nullptr, \
                     0, 0, qwo, beta_r)
  if (xf) {
    // stored-q schedule variants: FP64, 32x8 and 16x16 tiles, no RHS or tangent-linear mode
    constexpr bool kHave = sizeof(S) == 8 && ((TI == 32 && TJ == 8) || (TI == 16 && TJ == 16));
    if constexpr (kHave) {
      if (rhs_mode || tl_dc2) return hipErrorInvalidValue;
      if (xf == kStoreNoQ && kind == 0)
        SOLAR_STAGE_LAUNCH(0, 0, kStoreNoQ);
      else if (xf == kStoreNoQ && kind == 1)
        SOLAR_STAGE_LAUNCH(1, 0, kStoreNoQ);
      else if (xf == kStageRecon && kind == 1 && qwo)
        SOLAR_STAGE_LAUNCH(1, 0, kStageRecon);
      else
        return hipErrorInvalidValue;
      return hipGetLastError();
    } else {
      return hipErrorInvalidValue;
    }
  }
  if (tl_dc2) {
    if constexpr (sizeof(S) == 8 && !PART) {
      if (kind == 0)
        SOLAR_STAGE_LAUNCH(0, 2, 0);
      else if (kind == 1)
        SOLAR_STAGE_LAUNCH(1, 2, 0);
      else
        SOLAR_STAGE_LAUNCH(2, 2, 0);
    } else {
      return hipErrorInvalidValue;  // the tangent-linear stage is FP64 only, on tiles that divide n
    }
  } else if (kind == 0) {
    SOLAR_STAGE_LAUNCH(0, 0, 0);
  } else if (kind == 1) {
    SOLAR_STAGE_LAUNCH(1, 0, 0);
  } else {
    SOLAR_STAGE_LAUNCH(2, 0, 0);
  }
#undef SOLAR_STAGE_LAUNCH
  return hipGetLastError();
}

// Tile shapes (TI x TJ) compiled for the stage and filter kernels; the
// default is the measured best for the widest TI dividing n.  A requested
// shape that does not divide n falls back to the default.  When no compiled
// shape divides n (n not a multiple of 16), the partial-tile kernels run with
// 32-wide or 16-wide tiles, whichever pads n less (ties to 32).
static bool fits(int n, int ti, int tj) { return ti > 0 && tj > 0 && n % ti == 0 && n % tj == 0; }
static int pad_pick_ti(int n) { return ((n + 31) / 32) * 32 <= ((n + 15) / 16) * 16 ? 32 : 16; }

template <typename S>
static hipError_t launch_stage_s(const DevTables& T, const S* qin, S* qout, S* qw, const double* gxp, const double* gyp,
                                 const double* gxv, const double* gyv, int stage, double dt, int rhs_mode,
                                 hipStream_t st, const LaunchShape& ls, const double* c2f, double* divout,
                                 int div32, const double* tl_dc2, const double* tl_div, S* ecols, int xf, S* qwo,
                                 double beta_r) {
#define SOLAR_STAGE_ARGS \
  T, qin, qout, qw, gxp, gyp, gxv, gyv, stage, dt, rhs_mode, st, ls, c2f, divout, div32, tl_dc2, tl_div, ecols, xf, qwo, beta_r
  if (T.n % 16 != 0) {  // partial tiles: 32x8 or 16x16
    if (pad_pick_ti(T.n) == 32) return launch_stage_t<32, 8, S, true>(SOLAR_STAGE_ARGS);
    return launch_stage_t<16, 16, S, true>(SOLAR_STAGE_ARGS);
  }
  int ti = ls.stage_ti, tj = ls.stage_tj;
  // measured best: 32x8 (64x8 within 2%), then 16x16; the stored-q schedule
  // variants exist for these two only
  if (!fits(T.n, ti, tj) || (xf && !((ti == 32 && tj == 8) || (ti == 16 && tj == 16)))) {
    ti = pick_ti(T.n);
    tj = ti == 64 ? 8 : 256 / ti;
  }
#define SOLAR_STAGE_CASE(A, B) \
  if (ti == A && tj == B) return launch_stage_t<A, B, S>(SOLAR_STAGE_ARGS);
  SOLAR_STAGE_CASE(64, 8)
  SOLAR_STAGE_CASE(64, 4)
  SOLAR_STAGE_CASE(32, 8)
  SOLAR_STAGE_CASE(16, 16)
  SOLAR_STAGE_CASE(32, 16)
  SOLAR_STAGE_CASE(32, 4)
#undef SOLAR_STAGE_CASE
  ti = pick_ti(T.n);
  if (ti == 64) return launch_stage_t<64, 8, S>(SOLAR_STAGE_ARGS);
  if (ti == 32) return launch_stage_t<32, 8, S>(SOLAR_STAGE_ARGS);
  return launch_stage_t<16, 16, S>(SOLAR_STAGE_ARGS);
#undef SOLAR_STAGE_ARGS
}

hipError_t launch_stage(const DevTables& T, const void* qin, void* qout, void* qw, bool f32, const double* gxp,
                        const double* gyp, const double* gxv, const double* gyv, int stage, double dt, int rhs_mode,
                        hipStream_t st, const LaunchShape& ls, const double* c2f, double* divout, int div32,
                        const double* tl_dc2, const double* tl_div, void* ecols, int xf, void* qwo, double beta_r) {
  if (f32) {
    if (xf) return hipErrorInvalidValue;  // the stored-q schedule variants are FP64
    return launch_stage_s(T, static_cast<const float*>(qin), static_cast<float*>(qout), static_cast<float*>(qw), gxp,
                          gyp, gxv, gyv, stage, dt, rhs_mode, st, ls, c2f, divout, div32, tl_dc2, tl_div,
                          static_cast<float*>(ecols), 0, static_cast<float*>(nullptr), 0.0);
  }
  return launch_stage_s(T, static_cast<const double*>(qin), static_cast<double*>(qout), static_cast<double*>(qw), gxp,
                        gyp, gxv, gyv, stage, dt, rhs_mode, st, ls, c2f, divout, div32, tl_dc2, tl_div,
                        static_cast<double*>(ecols), xf, static_cast<double*>(qwo), beta_r);
}

template <int TI, int TJ, typename S, bool PART = false, bool PAIR = (sizeof(S) == 4)>
static hipError_t launch_filter_t(const DevTables& T, const S* qin, S* qout, const double* fgx, const double* fgy,
                                  hipStream_t st, const LaunchShape& ls) {
  const int tiles = ((T.n + TI - 1) / TI) * ((T.n + TJ - 1) / TJ) * 30;
  // radial chunks of >= 64 planes (each re-reads 10 halo planes), ~4k blocks: a single chunk per
  // column at 6x256^2x384 leaves a long tail (3.45 ms against 2.96 ms with two)
  const int nkch = ls.filter_chunks > 0 ? ls.filter_chunks : chunks_for(tiles, T.nr, 4096, 64);
  const int kchunk = (T.nr + nkch - 1) / nkch;
  dim3 grid(tiles * nkch), block(TI, TJ);
  hipLaunchKernelGGL((k_filter<TI, TJ, S, PART, PAIR>), grid, block, 0, st, T, qin, qout, fgx, fgy, kchunk, nkch);
  return hipGetLastError();
}

template <typename S>
static hipError_t launch_filter_s(const DevTables& T, const S* qin, S* qout, const double* fgx, const double* fgy,
                                  hipStream_t st, const LaunchShape& ls) {
  // FILTER_XI in paired tasks: always for an FP32 state, for FP64 on panels up to 128
  const bool pair = sizeof(S) == 4 || T.n <= 128;
  if (T.n % 16 != 0) {  // partial tiles
    if (pad_pick_ti(T.n) == 32)
      return pair ? launch_filter_t<32, 16, S, true, true>(T, qin, qout, fgx, fgy, st, ls)
                  : launch_filter_t<32, 16, S, true, false>(T, qin, qout, fgx, fgy, st, ls);
    return pair ? launch_filter_t<16, 16, S, true, true>(T, qin, qout, fgx, fgy, st, ls)
                : launch_filter_t<16, 16, S, true, false>(T, qin, qout, fgx, fgy, st, ls);
  }
  int ti = ls.filter_ti, tj = ls.filter_tj;
  if (!fits(T.n, ti, tj)) {  // measured best: 32x16 (n%32==0), 16x16
    ti = T.n % 32 == 0 ? 32 : 16;
    tj = 16;
  }
  if (pair && ti == 32 && tj == 16) return launch_filter_t<32, 16, S, false, true>(T, qin, qout, fgx, fgy, st, ls);
  if (pair && ti == 16 && tj == 16) return launch_filter_t<16, 16, S, false, true>(T, qin, qout, fgx, fgy, st, ls);
#define SOLAR_FILTER_CASE(A, B) \
  if (ti == A && tj == B) return launch_filter_t<A, B, S>(T, qin, qout, fgx, fgy, st, ls);
  SOLAR_FILTER_CASE(32, 16)
  SOLAR_FILTER_CASE(16, 16)
  SOLAR_FILTER_CASE(64, 4)
  SOLAR_FILTER_CASE(32, 8)
  SOLAR_FILTER_CASE(64, 8)
  SOLAR_FILTER_CASE(64, 16)
#undef SOLAR_FILTER_CASE
  if (T.n % 32 == 0)
    return pair ? launch_filter_t<32, 16, S, false, true>(T, qin, qout, fgx, fgy, st, ls)
                : launch_filter_t<32, 16, S>(T, qin, qout, fgx, fgy, st, ls);
  return pair ? launch_filter_t<16, 16, S, false, true>(T, qin, qout, fgx, fgy, st, ls)
              : launch_filter_t<16, 16, S>(T, qin, qout, fgx, fgy, st, ls);
}

hipError_t launch_filter(const DevTables& T, const void* qin, void* qout, bool f32, const double* fgx,
                         const double* fgy, hipStream_t st, const LaunchShape& ls) {
  if (f32)
    return launch_filter_s(T, static_cast<const float*>(qin), static_cast<float*>(qout), fgx, fgy, st, ls);
  return launch_filter_s(T, static_cast<const double*>(qin), static_cast<double*>(qout), fgx, fgy, st, ls);
}

hipError_t launch_scan(const void* q, bool f32, size_t nf, double* acc, int mode, hipStream_t st, int field) {
  if (field < 0 || field > 4) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, st, acc, 2);
  const size_t total = mode == 0 ? nf * 30 : nf * 6;
  if (f32)
    hipLaunchKernelGGL(k_scan<float>, dim3(1024), dim3(256), 0, st, static_cast<const float*>(q), total, nf, acc, mode,
                       field);
  else
    hipLaunchKernelGGL(k_scan<double>, dim3(1024), dim3(256), 0, st, static_cast<const double*>(q), total, nf, acc,
                       mode, field);
  return hipGetLastError();
}

hipError_t launch_place(const double* src, void* dst, bool f32, size_t nf, int dfield, hipStream_t st) {
  if (f32)
    hipLaunchKernelGGL(k_place<float>, dim3(1024), dim3(256), 0, st, src, static_cast<float*>(dst), nf, dfield);
  else
    hipLaunchKernelGGL(k_place<double>, dim3(1024), dim3(256), 0, st, src, static_cast<double*>(dst), nf, dfield);
  return hipGetLastError();
}

// ---------------------------------------------------------------- ensembles

template <int TI, int TJ>
static hipError_t launch_stage_ens_t(const EnsMember* M, const int4* items, int nitems, int grid, int stage, int cur,
                                     hipStream_t st) {
  const double al = kAlphaRk4[stage], be = kBetaRk4[stage];
  DevTables none{};
  dim3 g(grid), block(TI, TJ);
  if (stage == 0)
    hipLaunchKernelGGL((k_stage<TI, TJ, 0, double, 1>), g, block, 0, st, none, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr, nullptr, al, be, 0.0, 0, 0, 0, nullptr, nullptr, 0, nullptr, nullptr, nullptr, M, items,
                       nitems, cur, nullptr, 0.0);
  else if (stage < 4)
    hipLaunchKernelGGL((k_stage<TI, TJ, 1, double, 1>), g, block, 0, st, none, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr, nullptr, al, be, 0.0, 0, 0, 0, nullptr, nullptr, 0, nullptr, nullptr, nullptr, M, items,
                       nitems, cur, nullptr, 0.0);
  else
    hipLaunchKernelGGL((k_stage<TI, TJ, 2, double, 1>), g, block, 0, st, none, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr, nullptr, al, be, 0.0, 0, 0, 0, nullptr, nullptr, 0, nullptr, nullptr, nullptr, M, items,
                       nitems, cur, nullptr, 0.0);
  return hipGetLastError();
}

hipError_t launch_stage_ens(const EnsMember* M, const int* items, int nitems, int grid, int ti, int tj, int stage,
                            int cur, hipStream_t st) {
  const int4* it = reinterpret_cast<const int4*>(items);
  if (nitems <= 0) return hipSuccess;
  if (stage < 0 || stage > 4) return hipErrorInvalidValue;
  if (ti == 64 && tj == 8) return launch_stage_ens_t<64, 8>(M, it, nitems, grid, stage, cur, st);
  if (ti == 32 && tj == 8) return launch_stage_ens_t<32, 8>(M, it, nitems, grid, stage, cur, st);
  if (ti == 16 && tj == 16) return launch_stage_ens_t<16, 16>(M, it, nitems, grid, stage, cur, st);
  return hipErrorInvalidValue;
}

hipError_t launch_filter_ens(const EnsMember* M, const int* items, int nitems, int grid, int ti, int tj, int cur,
                             hipStream_t st) {
  const int4* it = reinterpret_cast<const int4*>(items);
  if (nitems <= 0) return hipSuccess;
  if (ti == 32 && tj == 16)
    hipLaunchKernelGGL((k_filter_ens<32, 16>), dim3(grid), dim3(32, 16), 0, st, M, it, nitems, cur);
  else if (ti == 16 && tj == 16)
    hipLaunchKernelGGL((k_filter_ens<16, 16>), dim3(grid), dim3(16, 16), 0, st, M, it, nitems, cur);
  else if (ti == 32 && tj == 8)
    hipLaunchKernelGGL((k_filter_ens<32, 8>), dim3(grid), dim3(32, 8), 0, st, M, it, nitems, cur);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template <bool VEC, int NI>
static hipError_t launch_gather_ens_v(const EnsMember* M, const int* table, int nmem, int nblocks, int maxn, int cur,
                                      int cols, int chunks, hipStream_t st) {
  const int nt = gather_threads(maxn, NI);
  const size_t lds = sizeof(double) * 2 * 3 * (VEC ? 3 : 1) * (size_t)maxn;
  hipError_t e = gather_lds(&k_gather_ens<VEC, NI>, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_gather_ens<VEC, NI>), dim3(nblocks, VEC ? 1 : 5), dim3(nt), lds, st, M, table, nmem, cur,
                     cols, chunks);
  return hipGetLastError();
}

hipError_t launch_gather_ens(const EnsMember* M, const int* table, int nmem, int nblocks, int maxn, bool vec, int cur,
                             bool cols, int chunks, hipStream_t st) {
  const bool two = 3 * maxn > 1024;  // n > 341: two strip points per thread
  const int c = cols ? 1 : 0;
  if (chunks < 1) return hipErrorInvalidValue;
  if (vec)
    return two ? launch_gather_ens_v<true, 2>(M, table, nmem, nblocks, maxn, cur, c, chunks, st)
               : launch_gather_ens_v<true, 1>(M, table, nmem, nblocks, maxn, cur, c, chunks, st);
  return two ? launch_gather_ens_v<false, 2>(M, table, nmem, nblocks, maxn, cur, c, chunks, st)
             : launch_gather_ens_v<false, 1>(M, table, nmem, nblocks, maxn, cur, c, chunks, st);
}

}  // namespace solar
