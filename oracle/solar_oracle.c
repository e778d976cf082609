/*
 * solar_oracle.c -- CPU restatement of the acoustic RK4 hot path of the
 * cubed-sphere solver (ankitvb/solar), used ONLY as the checker of the HIP
 * implementation in solar_amd/.  TEST INFRASTRUCTURE: nothing in the product
 * path links, loads or calls this file; tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg are its only users.
 *
 * It follows the reference loop for loop, for the single-rank-per-panel
 * decomposition px=py=pr=1 (6 panels, each an N x N x R block, i fastest):
 *   header:40-55,68-76,99-105,123-133   scheme constants (verbatim decimals)
 *   metric.f90:14-91                   grids, metric terms, ghost targets
 *   interpolation.f90:9-54,140-378     Lagrange weights, INTERPOLATE_ALONG_X/Y
 *   transform.f90:8-59,114-276         basis conversion + receiver-basis transform
 *   derivative.f90:8-758               DDX, DDY, DDR
 *   grad.f90:13-136                    COMPUTE_GRADIENT, COMPUTE_DIVERGENCE
 *   filter.f90:9-781                   FILTER_EXP, FILTER_XI/ETA/R
 *   compute_rhs_acoustic.f90:18-149    RHS (time-dependent forcing omitted)
 *   bc_acoustic_wave.f90:14-34, sponge.f90:20-50, timestep.f90:61-82
 * The MPI panel exchange (derivative.f90:55-133,374-452; filter.f90:93-171,
 * 365-443; partner ranks init_mpi.f90:164-304) becomes a copy of the sender
 * panel's strip argument into the receiver's ghost array, see route_x/route_y.
 *
 * Arithmetic order: every expression is evaluated in the order the Fortran
 * source writes it (left to right, products before sums), every stencil sum
 * starts from 0 and adds terms in ascending tap order.  Build with
 * -ffp-contract=off so no fused multiply-add changes the rounding.
 *
 * Two deliberate departures, both documented where they occur:
 *   - FILTER_XI (filter.f90:257-279) and FILTER_ETA (filter.f90:529-593)
 *     reverse the received ghost strips with loops that run past the strip
 *     array bounds (i to mx on an (nst,my,mr) array; j to my on an
 *     (mx,nst,mr) array).  The in-bounds intent is the same reversal that
 *     DDX/DDY perform (derivative.f90:217-282,538-561); that is what is
 *     implemented here.
 *   - The forcing term (compute_rhs_acoustic.f90:92-104,119-125) is omitted.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NST 3   /* header:23-25  noverlap = nst = 3 */
#define NORD 4  /* header:24 */
#define NSTR 5  /* header:26 */

/* header:40-46 DRP first-derivative coefficients (Tam & Webb). */
static const double coeffs1[7] = {-0.02651995, 0.18941314, -0.79926643, 0.0,
                                  0.79926643, -0.18941314, 0.02651995};
/* header:49-55 bdy_coeffs1(nst,nstencil), DATA fills column-major:
 * bdy[k][iter] = bdy_coeffs1(k+1, iter+1). */
static const double bdy_coeffs1[3][7] = {
    {-2.192280339, 4.748611401, -5.108851915, 4.461567104, -2.833498741, 1.128328861, -0.203876371},
    {-0.209337622, -1.084875676, 2.147776050, -1.388928322, 0.768949766, -0.28181465, 0.048230454},
    {0.049041958, -0.468840357, -0.474760914, 1.273274737, -0.518484526, 0.166138533, -0.026369431}};
/* header:68-76 low-storage RK(5,4) */
static const double alpha_rk4[5] = {0.0, -4.801594388477952e-01, -1.404247195200064e+00,
                                    -2.016477077503357e+00, -1.056444269767207e+00};
static const double beta_rk4[5] = {1.028639988104959e-01, 7.408540575766301e-01, 7.426530946683799e-01,
                                   4.694937902357905e-01, 1.881733382887932e-01};
/* header:99-105 */
static const double filt_coeffs[7] = {0.015625, -0.09375, 0.234375, 0.6875, 0.234375, -0.09375, 0.015625};
/* header:123-133 (truncated decimals, sum 0.999998) */
static const double filt_coeffs_r[11] = {0.0009765, -0.0097655, 0.043945, -0.1171875, 0.2050775, 0.753906,
                                         0.2050775, -0.1171875, 0.043945, -0.0097655, 0.0009765};

typedef struct {
  int N, R;
  int sponge_len;
  double a_sponge;
  double pi, dxi, deta, dchi;
  double *gxi, *geta, *tanxi, *taneta, *mC, *mD; /* N */
  double *gr, *gchi, *drdchi;                    /* R */
  double *mdelta, *sqrtd, *XY, *CD;              /* N*N, i fastest */
  double *interpoints;                           /* N*3 : [l*N + i] */
  double *interpolant;                           /* 9*N*3 : [(l*N + i)*9 + iter+4] */
  double *c2, *rho0, *g0;                        /* R radial background profiles */
  double* c2f; /* optional c2(i,j,k) per panel, (N,N,R,6); NULL = radial c2 */
  /* transform.f90:138-155 ghost-coordinate tables */
  double Xm[4], Xp[4], Ym[4], Yp[4], Cm[4], Cp[4], Dm[4], Dp[4];
  double *global_C, *global_D; /* N */
} orc_t;

#define IDX(N, i, j, k) ((size_t)(k) * (N) * (N) + (size_t)(j) * (N) + (i))

/* ------------------------------------------------------------------ setup */

/* metric.f90:14-91 (px=py=pr=1, so xi=gxi, eta=geta, chi=gchi, r=gr). */
static void setup_metrics(orc_t* o) {
  const int N = o->N, R = o->R;
  const double pi = o->pi;
  for (int i = 1; i <= N; i++) {
    o->gxi[i - 1] = -pi * 0.25 + (((double)i - 1.0) * pi * 0.5) / ((double)N - 1.0);
    o->geta[i - 1] = -pi * 0.25 + (((double)i - 1.0) * pi * 0.5) / ((double)N - 1.0);
  }
  for (int i = 1; i <= R; i++) o->gchi[i - 1] = 0.0 + (((double)i - 1.0) * 1.0) / ((double)R - 1.0);
  if (R == 1) {
    /* 2-D: metric.f90:20,45 would read 0/0 and chi(2), which does not exist;
     * nothing radial is used (DDR returns 0, derivative.f90:620-623), so the
     * radial metric is left at 0. */
    o->gchi[0] = 0.0;
    o->dchi = 0.0;
    o->drdchi[0] = 0.0;
  }
  for (int i = 0; i < N; i++) {
    o->tanxi[i] = tan(o->gxi[i]);
    o->taneta[i] = tan(o->geta[i]);
  }
  o->dxi = o->gxi[1] - o->gxi[0];
  o->deta = o->dxi;
  if (R == 1) goto angular;
  o->dchi = o->gchi[1] - o->gchi[0];

  /* metric.f90:48-67 gdrdchi */
  for (int kk = NST + 1; kk <= R - NST; kk++) {
    double s = 0.0;
    for (int it = -NST; it <= NST; it++) s = s + ((1.0 / o->dchi) * coeffs1[it + 3]) * o->gr[kk + it - 1];
    o->drdchi[kk - 1] = s;
  }
  for (int kk = 1; kk <= NST; kk++) {
    double s = 0.0;
    for (int it = 1; it <= 7; it++) s = s + ((1.0 / o->dchi) * bdy_coeffs1[kk - 1][it - 1]) * o->gr[it - 1];
    o->drdchi[kk - 1] = s;
  }
  for (int kk = R - NST + 1; kk <= R; kk++) {
    double s = 0.0;
    for (int it = 1; it <= 7; it++)
      s = s - ((1.0 / o->dchi) * bdy_coeffs1[R - kk][7 - it]) * o->gr[R - 7 + it - 1];
    o->drdchi[kk - 1] = s;
  }

angular:
  /* metric.f90:74-85 */
  for (int i = 0; i < N; i++) {
    o->mC[i] = pow(1.0 + o->tanxi[i] * o->tanxi[i], 0.5);
    o->mD[i] = pow(1.0 + o->taneta[i] * o->taneta[i], 0.5);
  }
  for (int j = 0; j < N; j++)
    for (int i = 0; i < N; i++) {
      double d = (1.0 + o->tanxi[i] * o->tanxi[i]) + o->taneta[j] * o->taneta[j];
      o->mdelta[j * N + i] = d;
      o->sqrtd[j * N + i] = pow(d, 0.5);
      o->XY[j * N + i] = o->tanxi[i] * o->taneta[j];
      o->CD[j * N + i] = o->mC[i] * o->mD[j];
    }
  /* metric.f90:89-91 */
  for (int l = 1; l <= NST; l++)
    for (int i = 0; i < N; i++)
      o->interpoints[(l - 1) * N + i] = atan(tan(o->geta[i]) / tan(o->gxi[N - 1] + l * o->dxi));

  /* transform.f90:142-155 */
  for (int i = 1; i <= NST; i++) {
    o->Xm[i] = tan(o->gxi[0] - i * o->dxi);
    o->Ym[i] = tan(o->geta[0] - i * o->deta);
    o->Cm[i] = sqrt(1.0 + o->Xm[i] * o->Xm[i]);
    o->Dm[i] = sqrt(1.0 + o->Ym[i] * o->Ym[i]);
    o->Xp[i] = tan(o->gxi[N - 1] + i * o->dxi);
    o->Yp[i] = tan(o->geta[N - 1] + i * o->deta);
    o->Cp[i] = sqrt(1.0 + o->Xp[i] * o->Xp[i]);
    o->Dp[i] = sqrt(1.0 + o->Yp[i] * o->Yp[i]);
  }
  for (int i = 0; i < N; i++) {
    double tx = tan(o->gxi[i]), te = tan(o->geta[i]);
    o->global_C[i] = pow(1.0 + tx * tx, 0.5);
    o->global_D[i] = pow(1.0 + te * te, 0.5);
  }
}

/* interpolation.f90:9-54 SETUP_LAGRANGE_INTERPOLANT */
static void setup_interpolant(orc_t* o) {
  const int N = o->N;
  double* w = o->interpolant;
#define W(it, i, l) w[((size_t)((l)-1) * N + ((i)-1)) * 9 + (it) + 4]
  memset(w, 0, sizeof(double) * 9 * N * 3);
  for (int k = 1; k <= NST; k++) {
    W(k, 1, k) = 1.0;
    W(-k, N, k) = 1.0;
    for (int i = 2; i <= N - 1; i++) {
      int lower = -NORD, upper = NORD;
      if (i < NORD + 1) lower = -i + 1;
      if (i > N - NORD) upper = N - i;
      for (int it = lower; it <= upper; it++) W(it, i, k) = 1.0;
      const double x = o->interpoints[(k - 1) * N + (i - 1)];
      for (int ip = lower; ip <= upper; ip++)
        for (int it = lower; it <= upper; it++)
          if (it != ip)
            W(ip, i, k) = ((x - o->geta[i + it - 1]) * W(ip, i, k)) / (o->geta[i + ip - 1] - o->geta[i + it - 1]);
    }
  }
#undef W
}

orc_t* orc_create(int N, int R, const double* gr, const double* c2, const double* rho0, const double* g0,
                  int sponge_len, double a_sponge) {
  if (N < 9 || (R < 12 && R != 1)) return NULL;  /* R == 1: the reference's 2-D case */
  orc_t* o = (orc_t*)calloc(1, sizeof(orc_t));
  o->N = N;
  o->R = R;
  o->sponge_len = sponge_len;
  o->a_sponge = a_sponge;
  o->pi = acos(-1.0);
  double** n1[] = {&o->gxi, &o->geta, &o->tanxi, &o->taneta, &o->mC, &o->mD, &o->global_C, &o->global_D};
  for (size_t a = 0; a < sizeof(n1) / sizeof(n1[0]); a++) *n1[a] = (double*)calloc(N, sizeof(double));
  double** r1[] = {&o->gr, &o->gchi, &o->drdchi, &o->c2, &o->rho0, &o->g0};
  for (size_t a = 0; a < sizeof(r1) / sizeof(r1[0]); a++) *r1[a] = (double*)calloc(R, sizeof(double));
  double** n2[] = {&o->mdelta, &o->sqrtd, &o->XY, &o->CD};
  for (size_t a = 0; a < sizeof(n2) / sizeof(n2[0]); a++) *n2[a] = (double*)calloc((size_t)N * N, sizeof(double));
  o->interpoints = (double*)calloc((size_t)N * 3, sizeof(double));
  o->interpolant = (double*)calloc((size_t)N * 27, sizeof(double));
  memcpy(o->gr, gr, sizeof(double) * R);
  if (R == 1) o->gr[0] = 1.0; /* metric.f90:35-38: the r effect is nullified in 2-D */
  memcpy(o->c2, c2, sizeof(double) * R);
  memcpy(o->rho0, rho0, sizeof(double) * R);
  memcpy(o->g0, g0, sizeof(double) * R);
  setup_metrics(o);
  setup_interpolant(o);
  return o;
}

void orc_destroy(orc_t* o) {
  if (!o) return;
  double* ps[] = {o->gxi, o->geta, o->tanxi, o->taneta, o->mC, o->mD, o->global_C, o->global_D, o->gr,
                  o->gchi, o->drdchi, o->c2, o->rho0, o->g0, o->mdelta, o->sqrtd, o->XY, o->CD,
                  o->interpoints, o->interpolant};
  for (size_t a = 0; a < sizeof(ps) / sizeof(ps[0]); a++) free(ps[a]);
  free(o->c2f);
  free(o);
}

/* The reference's c2 is the 3-D array c2(mx,my,mr) (compute_rhs_acoustic.f90:134,
 * ic_acoustic_wave.f90:94-108 fills it from a radial profile).  A non-NULL
 * field c2f(N,N,R,6) replaces the radial profile in COMPUTE_RHS. */
void orc_set_c2_field(orc_t* o, const double* c2f) {
  const size_t cnt = (size_t)6 * o->N * o->N * o->R;
  if (!c2f) {
    free(o->c2f);
    o->c2f = NULL;
    return;
  }
  if (!o->c2f) o->c2f = (double*)calloc(cnt, sizeof(double));
  memcpy(o->c2f, c2f, cnt * sizeof(double));
}

/* Table export for the table-parity test. */
void orc_get_tables(const orc_t* o, double* scal /*4: dxi,deta,dchi,pi*/, double* n1 /*8*N*/, double* r1 /*3*R*/,
                    double* n2 /*4*N*N*/, double* interp /*27*N*/, double* ghost_coords /*8*3*/) {
  const int N = o->N, R = o->R;
  scal[0] = o->dxi; scal[1] = o->deta; scal[2] = o->dchi; scal[3] = o->pi;
  const double* a1[] = {o->gxi, o->geta, o->tanxi, o->taneta, o->mC, o->mD, o->global_C, o->global_D};
  for (int a = 0; a < 8; a++) memcpy(n1 + (size_t)a * N, a1[a], sizeof(double) * N);
  memcpy(r1, o->gchi, sizeof(double) * R);
  memcpy(r1 + R, o->drdchi, sizeof(double) * R);
  memcpy(r1 + 2 * R, o->gr, sizeof(double) * R);
  const double* a2[] = {o->mdelta, o->sqrtd, o->XY, o->CD};
  for (int a = 0; a < 4; a++) memcpy(n2 + (size_t)a * N * N, a2[a], sizeof(double) * N * N);
  memcpy(interp, o->interpolant, sizeof(double) * 27 * N);
  const double* g[8] = {o->Xm, o->Xp, o->Ym, o->Yp, o->Cm, o->Cp, o->Dm, o->Dp};
  for (int a = 0; a < 8; a++)
    for (int l = 1; l <= 3; l++) ghost_coords[a * 3 + l - 1] = g[a][l];
}

/* ------------------------------------------------------- interpolation */

/* interpolation.f90:261-378 INTERPOLATE_ALONG_X, xrank=0=px-1.
 * in/out: (N,3,R) [(k*3 + l)*N + i]. */
static void interp_x(const orc_t* o, const double* in, double* out) {
  const int N = o->N, R = o->R;
#define IN(i, l, k) in[((size_t)(k) * 3 + (l)) * N + (i)-1]
#define OUT(i, l, k) out[((size_t)(k) * 3 + (l)) * N + (i)-1]
#define WT(it, i, l) o->interpolant[((size_t)(l) * N + ((i)-1)) * 9 + (it) + 4]
#pragma omp parallel for schedule(static)
  for (int k = 0; k < R; k++)
    for (int l = 0; l < 3; l++) {
      const int j = l + 1; /* Fortran layer index */
      OUT(1, l, k) = IN(j + 1, l, k);
      OUT(N, l, k) = IN(N - j, l, k);
      for (int i = NORD + 1; i <= N - NORD; i++) {
        double s = 0.0;
        for (int it = -NORD; it <= NORD; it++) s = s + IN(i + it, l, k) * WT(it, i, l);
        OUT(i, l, k) = s;
      }
      for (int i = 2; i <= NORD; i++) {
        double s = 0.0;
        for (int it = -i + 1; it <= NORD; it++) s = s + IN(it + i, l, k) * WT(it, i, l);
        OUT(i, l, k) = s;
      }
      for (int i = N - NORD + 1; i <= N - 1; i++) {
        double s = 0.0;
        for (int it = -NORD; it <= N - i; it++) s = s + IN(i + it, l, k) * WT(it, i, l);
        OUT(i, l, k) = s;
      }
    }
#undef IN
#undef OUT
}

/* interpolation.f90:140-257 INTERPOLATE_ALONG_Y, yrank=0=py-1.
 * in/out: (3,N,R) [(k*N + j)*3 + l]. */
static void interp_y(const orc_t* o, const double* in, double* out) {
  const int N = o->N, R = o->R;
#define IN(l, j, k) in[((size_t)(k) * N + (j)-1) * 3 + (l)]
#define OUT(l, j, k) out[((size_t)(k) * N + (j)-1) * 3 + (l)]
#pragma omp parallel for schedule(static)
  for (int k = 0; k < R; k++)
    for (int l = 0; l < 3; l++) {
      const int i = l + 1;
      OUT(l, 1, k) = IN(l, i + 1, k);
      OUT(l, N, k) = IN(l, N - i, k);
      for (int j = NORD + 1; j <= N - NORD; j++) {
        double s = 0.0;
        for (int it = -NORD; it <= NORD; it++) s = s + IN(l, j + it, k) * WT(it, j, l);
        OUT(l, j, k) = s;
      }
      for (int j = 2; j <= NORD; j++) {
        double s = 0.0;
        for (int it = -j + 1; it <= NORD; it++) s = s + IN(l, j + it, k) * WT(it, j, l);
        OUT(l, j, k) = s;
      }
      for (int j = N - NORD + 1; j <= N - 1; j++) {
        double s = 0.0;
        for (int it = -NORD; it <= N - j; it++) s = s + IN(l, j + it, k) * WT(it, j, l);
        OUT(l, j, k) = s;
      }
    }
#undef IN
#undef OUT
#undef WT
}

/* Strip extraction: a(:,2:4,:), a(:,my-1:my-3:-1,:), a(2:4,:,:), a(mx-1:mx-3:-1,:,:)
 * (grad.f90:28-31).  south/north -> (N,3,R), west/east -> (3,N,R). */
static void extract_strips(const orc_t* o, const double* a, double* s, double* n, double* w, double* e) {
  const int N = o->N, R = o->R;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < R; k++)
    for (int l = 0; l < 3; l++) {
      for (int i = 0; i < N; i++) {
        s[((size_t)k * 3 + l) * N + i] = a[IDX(N, i, 1 + l, k)];
        n[((size_t)k * 3 + l) * N + i] = a[IDX(N, i, N - 2 - l, k)];
      }
      for (int j = 0; j < N; j++) {
        w[((size_t)k * N + j) * 3 + l] = a[IDX(N, 1 + l, j, k)];
        e[((size_t)k * N + j) * 3 + l] = a[IDX(N, N - 2 - l, j, k)];
      }
    }
}

/* (N,3,R) <-> (3,N,R) transposes of grad.f90:34-43 */
static void transpose_xy(const orc_t* o, const double* in_N3R, double* out_3NR) {
  const int N = o->N, R = o->R;
  for (int k = 0; k < R; k++)
    for (int l = 0; l < 3; l++)
      for (int i = 0; i < N; i++) out_3NR[((size_t)k * N + i) * 3 + l] = in_N3R[((size_t)k * 3 + l) * N + i];
}
static void transpose_yx(const orc_t* o, const double* in_3NR, double* out_N3R) {
  const int N = o->N, R = o->R;
  for (int k = 0; k < R; k++)
    for (int l = 0; l < 3; l++)
      for (int i = 0; i < N; i++) out_N3R[((size_t)k * 3 + l) * N + i] = in_3NR[((size_t)k * N + i) * 3 + l];
}

/* ------------------------------------------------------------ routing */

/* Arguments every panel passes to a DDX/FILTER_XI call (each (3,N,R)) and to
 * a DDY/FILTER_ETA call (each (N,3,R)). */
enum { ARG_LOW = 0, ARG_HIGH = 1, ARG_LEFT = 2, ARG_RIGHT = 3 };
typedef struct {
  int sender; /* 0-based panel */
  int arg;    /* which strip argument of the sender's call */
  int rev;    /* along-edge reversal applied by the receiver */
} route_t;

/* DDX receives (derivative.f90:374-452 sends/receives by intercommunicator,
 * partners init_mpi.f90:164-304; reversal derivative.f90:538-561).
 * [receiver][0] = fp (i<1 ghost), [receiver][1] = fn (i>N ghost). */
static const route_t route_x[6][2] = {
    {{3, ARG_RIGHT, 0}, {1, ARG_LEFT, 0}}, /* 1: fp <- 4.right, fn <- 2.left */
    {{0, ARG_RIGHT, 0}, {2, ARG_LEFT, 0}}, /* 2 */
    {{1, ARG_RIGHT, 0}, {3, ARG_LEFT, 0}}, /* 3 */
    {{2, ARG_RIGHT, 0}, {0, ARG_LEFT, 0}}, /* 4 */
    {{3, ARG_HIGH, 1}, {1, ARG_HIGH, 0}},  /* 5: fp <- 4.high rev j, fn <- 2.high */
    {{3, ARG_LOW, 0}, {1, ARG_LOW, 1}},    /* 6: fp <- 4.low, fn <- 2.low rev j */
};
/* DDY receives (derivative.f90:55-133; reversal derivative.f90:217-282). */
static const route_t route_y[6][2] = {
    {{5, ARG_HIGH, 0}, {4, ARG_LOW, 0}},   /* 1: fp <- 6.high, fn <- 5.low */
    {{5, ARG_RIGHT, 1}, {4, ARG_RIGHT, 0}},/* 2: fp <- 6.right rev i, fn <- 5.right */
    {{5, ARG_LOW, 1}, {4, ARG_HIGH, 1}},   /* 3: fp <- 6.low rev i, fn <- 5.high rev i */
    {{5, ARG_LEFT, 0}, {4, ARG_LEFT, 1}},  /* 4: fp <- 6.left, fn <- 5.left rev i */
    {{0, ARG_HIGH, 0}, {2, ARG_HIGH, 1}},  /* 5: fp <- 1.high, fn <- 3.high rev i */
    {{2, ARG_LOW, 1}, {0, ARG_LOW, 0}},    /* 6: fp <- 3.low rev i, fn <- 1.low */
};

/* Receive an x-ghost (3,N,R) with optional j reversal. */
static void recv_x(const orc_t* o, const double* src, double* dst, int rev) {
  const int N = o->N, R = o->R;
  for (int k = 0; k < R; k++)
    for (int j = 0; j < N; j++)
      for (int l = 0; l < 3; l++)
        dst[((size_t)k * N + j) * 3 + l] = src[((size_t)k * N + (rev ? N - 1 - j : j)) * 3 + l];
}
/* Receive a y-ghost (N,3,R) with optional i reversal. */
static void recv_y(const orc_t* o, const double* src, double* dst, int rev) {
  const int N = o->N, R = o->R;
  for (int k = 0; k < R; k++)
    for (int l = 0; l < 3; l++)
      for (int i = 0; i < N; i++)
        dst[((size_t)k * 3 + l) * N + i] = src[((size_t)k * 3 + l) * N + (rev ? N - 1 - i : i)];
}

/* ----------------------------------------------------------- derivatives */

/* derivative.f90:327-601 DDX with px=1.  fp, fn: (3,N,R). */
static void ddx(const orc_t* o, const double* f, double* d, const double* fp, const double* fn) {
  const int N = o->N, R = o->R;
  const double fac = 1.0 / o->dxi;
#define FP(l, j, k) fp[((size_t)(k) * N + (j)) * 3 + (l)-1]
#define FN(l, j, k) fn[((size_t)(k) * N + (j)) * 3 + (l)-1]
#pragma omp parallel for schedule(static)
  for (int k = 0; k < R; k++)
    for (int j = 0; j < N; j++) {
      const double* fl = f + IDX(N, 0, j, k);
      double* dl = d + IDX(N, 0, j, k);
      for (int i = NST + 1; i <= N - NST; i++) {
        double s = 0.0;
        for (int it = -NST; it <= NST; it++) s = s + (fac * coeffs1[it + 3]) * fl[i + it - 1];
        dl[i - 1] = s;
      }
      for (int i = 1; i <= NST; i++) {
        double s = 0.0;
        for (int it = -NST; it <= -i; it++) s = s + (fac * coeffs1[it + 3]) * FP(-i - it + 1, j, k);
        for (int it = -i + 1; it <= NST; it++) s = s + (fac * coeffs1[it + 3]) * fl[i + it - 1];
        dl[i - 1] = s;
      }
      for (int i = N - NST + 1; i <= N; i++) {
        double s = 0.0;
        for (int it = -NST; it <= N - i; it++) s = s + (fac * coeffs1[it + 3]) * fl[i + it - 1];
        for (int it = N - i + 1; it <= NST; it++) s = s + (fac * coeffs1[it + 3]) * FN(i + it - N, j, k);
        dl[i - 1] = s;
      }
    }
#undef FP
#undef FN
}

/* derivative.f90:8-323 DDY with py=1.  fp, fn: (N,3,R). */
static void ddy(const orc_t* o, const double* f, double* d, const double* fp, const double* fn) {
  const int N = o->N, R = o->R;
  const double fac = 1.0 / o->deta;
#define FP(i, l, k) fp[((size_t)(k) * 3 + (l)-1) * N + (i)]
#define FN(i, l, k) fn[((size_t)(k) * 3 + (l)-1) * N + (i)]
#pragma omp parallel for schedule(static)
  for (int k = 0; k < R; k++) {
    for (int j = NST + 1; j <= N - NST; j++)
      for (int i = 0; i < N; i++) {
        double s = 0.0;
        for (int it = -NST; it <= NST; it++) s = s + (fac * coeffs1[it + 3]) * f[IDX(N, i, j + it - 1, k)];
        d[IDX(N, i, j - 1, k)] = s;
      }
    for (int j = 1; j <= NST; j++)
      for (int i = 0; i < N; i++) {
        double s = 0.0;
        for (int it = -NST; it <= -j; it++) s = s + (fac * coeffs1[it + 3]) * FP(i, -it - j + 1, k);
        for (int it = -j + 1; it <= NST; it++) s = s + (fac * coeffs1[it + 3]) * f[IDX(N, i, it + j - 1, k)];
        d[IDX(N, i, j - 1, k)] = s;
      }
    for (int j = N - NST + 1; j <= N; j++)
      for (int i = 0; i < N; i++) {
        double s = 0.0;
        for (int it = -NST; it <= N - j; it++) s = s + (fac * coeffs1[it + 3]) * f[IDX(N, i, it + j - 1, k)];
        for (int it = N - j + 1; it <= NST; it++) s = s + (fac * coeffs1[it + 3]) * FN(i, it + j - N, k);
        d[IDX(N, i, j - 1, k)] = s;
      }
  }
#undef FP
#undef FN
}

/* derivative.f90:605-758 DDR with pr=1 (rrank=0=pr-1): Tam-Dong closures at
 * both shell ends.  f, d: (nx,ny,R) with nx*ny = ncol columns. */
static void ddr_cols(const orc_t* o, const double* f, double* d, size_t ncol) {
  const int R = o->R;
  if (R == 1) { /* 2-D case, derivative.f90:620-623: dr_f = 0 */
    for (size_t c = 0; c < ncol; c++) d[c] = 0.0;
    return;
  }
  const double fac = 1.0 / o->dchi;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= R; k++) {
    double* dk = d + (size_t)(k - 1) * ncol;
    if (k >= NST + 1 && k <= R - NST) {
      const double fac2 = fac / o->drdchi[k - 1];
      for (size_t c = 0; c < ncol; c++) {
        double s = 0.0;
        for (int it = -NST; it <= NST; it++) s = s + (fac2 * coeffs1[it + 3]) * f[(size_t)(k + it - 1) * ncol + c];
        dk[c] = s;
      }
    } else if (k <= NST) {
      const double fac2 = fac / o->drdchi[k - 1];
      for (size_t c = 0; c < ncol; c++) {
        double s = 0.0;
        for (int it = 1; it <= 7; it++) s = s + (fac2 * bdy_coeffs1[k - 1][it - 1]) * f[(size_t)(it - 1) * ncol + c];
        dk[c] = s;
      }
    } else {
      const double fac2 = -fac / o->drdchi[k - 1];
      for (size_t c = 0; c < ncol; c++) {
        double s = 0.0;
        for (int it = 1; it <= 7; it++)
          s = s + (fac2 * bdy_coeffs1[R - k][7 - it]) * f[(size_t)(R - 7 + it - 1) * ncol + c];
        dk[c] = s;
      }
    }
  }
}

/* ----------------------------------------------------------- transforms */

/* transform.f90:8-59 CONVERT_XI_ETA_TO_THETA_PHI for panel z (1-based). */
static void convert_to_theta_phi(const orc_t* o, int z, const double* axi, const double* aeta, double* ath,
                                 double* aph) {
  const int N = o->N, R = o->R;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < R; k++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        const size_t p = IDX(N, i, j, k);
        const int ij = j * N + i;
        if (z >= 1 && z <= 4) {
          ath[p] = (o->XY[ij] / o->CD[ij]) * axi[p] - aeta[p];
          aph[p] = (o->sqrtd[ij] / o->CD[ij]) * axi[p];
        } else if (z == 5) {
          const double factor = 1.0 / pow(o->mdelta[ij] - 1.0, 0.5);
          ath[p] = factor * ((o->tanxi[i] / o->mD[j]) * axi[p] + (o->taneta[j] / o->mC[i]) * aeta[p]);
          aph[p] = factor * (-(((o->sqrtd[ij] * o->taneta[j]) / o->mD[j]) * axi[p]) +
                             ((o->sqrtd[ij] * o->tanxi[i]) / o->mC[i]) * aeta[p]);
        } else {
          const double factor = 1.0 / pow(o->mdelta[ij] - 1.0, 0.5);
          ath[p] = -(factor * ((o->tanxi[i] / o->mD[j]) * axi[p] + (o->taneta[j] / o->mC[i]) * aeta[p]));
          aph[p] = factor * (((o->sqrtd[ij] * o->taneta[j]) / o->mD[j]) * axi[p] -
                             ((o->sqrtd[ij] * o->tanxi[i]) / o->mC[i]) * aeta[p]);
        }
      }
}

/* Divergence strips of one sender panel after TRANSFORM_VECTOR_BETWEEN_MESHES_DIV
 * (transform.f90:161-274).  Inputs: interpolated theta/phi strips
 * (south/north (N,3,R), west/east (3,N,R)).  Outputs the receiver-basis
 * strips: *_xi for DDX and *_eta for DDY; entries a panel does not send are
 * left untouched. */
static void transform_div(const orc_t* o, int z, const double* sth, const double* sph, const double* nth,
                          const double* nph, const double* wth, const double* wph, const double* eth,
                          const double* eph, double* east_xi, double* west_xi, double* south_eta,
                          double* north_eta, double* east_eta, double* west_eta, double* north_xi,
                          double* south_xi) {
  const int N = o->N, R = o->R;
  const double *X_m = o->Xm, *X_p = o->Xp, *Y_m = o->Ym, *Y_p = o->Yp, *C_m = o->Cm, *C_p = o->Cp;
  const double* gD = o->global_D;
#define S3(i, l, k) (((size_t)(k) * 3 + (l)) * N + (i)) /* (N,3,R) */
#define T3(l, j, k) (((size_t)(k) * N + (j)) * 3 + (l)) /* (3,N,R) */
  if (z >= 1 && z <= 4) {
    for (int k = 0; k < R; k++) {
      for (int l = 1; l <= 3; l++)
        for (int j = 0; j < N; j++) {
          const double t = tan(o->geta[j]);
          east_xi[T3(l - 1, j, k)] = ((C_m[l] * gD[j]) / sqrt((1.0 + X_m[l] * X_m[l]) + t * t)) * eph[T3(l - 1, j, k)];
        }
      for (int l = 1; l <= 3; l++)
        for (int j = 0; j < N; j++) {
          const double t = tan(o->geta[j]);
          west_xi[T3(l - 1, j, k)] = ((C_p[l] * gD[j]) / sqrt((1.0 + X_p[l] * X_p[l]) + t * t)) * wph[T3(l - 1, j, k)];
        }
    }
  }
  if (z == 5 || z == 6) {
    const double* Y = (z == 5) ? Y_p : Y_m;
    for (int k = 0; k < R; k++)
      for (int l = 1; l <= 3; l++)
        for (int i = 0; i < N; i++) {
          const double tx = tan(o->gxi[i]); /* tan(xi(i)) and tan(gxi(ii)) coincide for px=py=1 */
          const double b = (tx * Y[l]) / sqrt((1.0 + tx * tx) + Y[l] * Y[l]);
          south_eta[S3(i, l - 1, k)] = -sth[S3(i, l - 1, k)] + b * sph[S3(i, l - 1, k)];
          east_eta[T3(l - 1, i, k)] = -eth[T3(l - 1, i, k)] + b * eph[T3(l - 1, i, k)];
          north_eta[S3(i, l - 1, k)] = -nth[S3(i, l - 1, k)] + b * nph[S3(i, l - 1, k)];
          west_eta[T3(l - 1, i, k)] = -wth[T3(l - 1, i, k)] + b * wph[T3(l - 1, i, k)];
        }
  }
  /* (1,5),(2,5),(3,5),(4,5): north strips, transform.f90:226-248 */
  for (int k = 0; k < R; k++)
    for (int l = 1; l <= 3; l++)
      for (int i = 0; i < N; i++) {
        const double tx = tan(o->gxi[i]);
        const double te = tan(o->geta[i]);
        const size_t s = S3(i, l - 1, k);
        if (z == 1 || z == 3) {
          const double Yv = (z == 1) ? Y_m[l] : Y_p[l];
          const double A = sqrt(tx * tx + Yv * Yv);
          const double B = sqrt((1.0 + tx * tx) + Yv * Yv);
          north_eta[s] = ((o->mC[i] * Yv) / A) * nth[s] + (((o->mC[i] / A) * tx) / B) * nph[s];
        }
        if (z == 2 || z == 4) {
          const double Xv = (z == 2) ? X_p[l] : X_m[l];
          const double A = sqrt(Xv * Xv + te * te);
          const double B = sqrt((1.0 + Xv * Xv) + te * te);
          north_xi[s] = ((gD[i] * Xv) / A) * nth[s] - (((gD[i] / A) * te) / B) * nph[s];
        }
      }
  /* (1,6),(2,6),(3,6),(4,6): south strips, transform.f90:252-274 */
  for (int k = 0; k < R; k++)
    for (int l = 1; l <= 3; l++)
      for (int i = 0; i < N; i++) {
        const double tx = tan(o->gxi[i]);
        const double te = tan(o->geta[i]);
        const size_t s = S3(i, l - 1, k);
        if (z == 1 || z == 3) {
          const double Yv = (z == 1) ? Y_p[l] : Y_m[l];
          const double A = sqrt(tx * tx + Yv * Yv);
          const double B = sqrt((1.0 + tx * tx) + Yv * Yv);
          south_eta[s] = -(((o->mC[i] * Yv) / A) * sth[s]) - ((((o->mC[i] / A) * tx) / B) * sph[s]);
        }
        if (z == 2 || z == 4) {
          const double Xv = (z == 2) ? X_p[l] : X_m[l];
          const double A = sqrt(Xv * Xv + te * te);
          const double B = sqrt((1.0 + Xv * Xv) + te * te);
          south_xi[s] = -(((gD[i] * Xv) / A) * sth[s]) + (((gD[i] / A) * te) / B) * sph[s];
        }
      }
#undef S3
#undef T3
}

/* -------------------------------------------------- per-panel workspace */

typedef struct {
  double* xarg[4]; /* DDX / FILTER_XI strip arguments: low, high, left, right, (3,N,R) */
  double* yarg[4]; /* DDY / FILTER_ETA strip arguments, (N,3,R) */
} panel_args_t;

static double* dalloc(size_t n) { return (double*)calloc(n, sizeof(double)); }

static void args_alloc(const orc_t* o, panel_args_t* pa) {
  const size_t ns = (size_t)3 * o->N * o->R;
  for (int p = 0; p < 6; p++)
    for (int a = 0; a < 4; a++) {
      pa[p].xarg[a] = dalloc(ns);
      pa[p].yarg[a] = dalloc(ns);
    }
}
static void args_free(panel_args_t* pa) {
  for (int p = 0; p < 6; p++)
    for (int a = 0; a < 4; a++) {
      free(pa[p].xarg[a]);
      free(pa[p].yarg[a]);
    }
}

/* Scalar strips of one panel field as its DDX/DDY (or FILTER_XI/ETA)
 * arguments: grad.f90:28-46, filter.f90:25-43. */
static void scalar_args(const orc_t* o, const double* a, panel_args_t* pa) {
  const size_t ns = (size_t)3 * o->N * o->R;
  double *s = dalloc(ns), *n = dalloc(ns), *w = dalloc(ns), *e = dalloc(ns);
  double *south = dalloc(ns), *north = dalloc(ns), *west = dalloc(ns), *east = dalloc(ns);
  extract_strips(o, a, s, n, w, e);
  interp_x(o, s, south);
  interp_x(o, n, north);
  interp_y(o, w, west);
  interp_y(o, e, east);
  transpose_xy(o, south, pa->xarg[ARG_LOW]);
  transpose_xy(o, north, pa->xarg[ARG_HIGH]);
  memcpy(pa->xarg[ARG_LEFT], west, ns * sizeof(double));
  memcpy(pa->xarg[ARG_RIGHT], east, ns * sizeof(double));
  memcpy(pa->yarg[ARG_LOW], south, ns * sizeof(double));
  memcpy(pa->yarg[ARG_HIGH], north, ns * sizeof(double));
  transpose_yx(o, west, pa->yarg[ARG_LEFT]);
  transpose_yx(o, east, pa->yarg[ARG_RIGHT]);
  free(s); free(n); free(w); free(e);
  free(south); free(north); free(west); free(east);
}

/* Divergence strips of one panel: grad.f90:95-119. */
static void vector_args(const orc_t* o, int z, const double* axi, const double* aeta, panel_args_t* pa) {
  const int N = o->N, R = o->R;
  const size_t ns = (size_t)3 * N * R, nf = (size_t)N * N * R;
  double *ath = dalloc(nf), *aph = dalloc(nf);
  convert_to_theta_phi(o, z, axi, aeta, ath, aph);
  double *s = dalloc(ns), *n = dalloc(ns), *w = dalloc(ns), *e = dalloc(ns);
  double* st[8];
  for (int a = 0; a < 8; a++) st[a] = dalloc(ns);
  /* st: 0 south_theta 1 north_theta 2 west_theta 3 east_theta 4..7 phi */
  extract_strips(o, ath, s, n, w, e);
  interp_x(o, s, st[0]);
  interp_x(o, n, st[1]);
  interp_y(o, w, st[2]);
  interp_y(o, e, st[3]);
  extract_strips(o, aph, s, n, w, e);
  interp_x(o, s, st[4]);
  interp_x(o, n, st[5]);
  interp_y(o, e, st[7]);
  interp_y(o, w, st[6]);
  double *east_xi = dalloc(ns), *west_xi = dalloc(ns), *south_eta = dalloc(ns), *north_eta = dalloc(ns);
  double *east_eta = dalloc(ns), *west_eta = dalloc(ns), *north_xi = dalloc(ns), *south_xi = dalloc(ns);
  transform_div(o, z, st[0], st[4], st[1], st[5], st[2], st[6], st[3], st[7], east_xi, west_xi, south_eta,
                north_eta, east_eta, west_eta, north_xi, south_xi);
  /* grad.f90:110-122 */
  transpose_xy(o, south_xi, pa->xarg[ARG_LOW]);
  transpose_xy(o, north_xi, pa->xarg[ARG_HIGH]);
  memcpy(pa->xarg[ARG_LEFT], west_xi, ns * sizeof(double));
  memcpy(pa->xarg[ARG_RIGHT], east_xi, ns * sizeof(double));
  memcpy(pa->yarg[ARG_LOW], south_eta, ns * sizeof(double));
  memcpy(pa->yarg[ARG_HIGH], north_eta, ns * sizeof(double));
  transpose_yx(o, west_eta, pa->yarg[ARG_LEFT]);
  transpose_yx(o, east_eta, pa->yarg[ARG_RIGHT]);
  free(ath); free(aph); free(s); free(n); free(w); free(e);
  for (int a = 0; a < 8; a++) free(st[a]);
  free(east_xi); free(west_xi); free(south_eta); free(north_eta);
  free(east_eta); free(west_eta); free(north_xi); free(south_xi);
}

/* Ghosts received by panel p (0-based) for its DDX and DDY calls. */
static void exchange(const orc_t* o, const panel_args_t* pa, int p, double* xfp, double* xfn, double* yfp,
                     double* yfn) {
  const route_t* rx = route_x[p];
  const route_t* ry = route_y[p];
  recv_x(o, pa[rx[0].sender].xarg[rx[0].arg], xfp, rx[0].rev);
  recv_x(o, pa[rx[1].sender].xarg[rx[1].arg], xfn, rx[1].rev);
  recv_y(o, pa[ry[0].sender].yarg[ry[0].arg], yfp, ry[0].rev);
  recv_y(o, pa[ry[1].sender].yarg[ry[1].arg], yfn, ry[1].rev);
}

/* ------------------------------------------------------- grad and div */

/* grad.f90:13-59 on all six panels. a, gx, ge, gr: 6 consecutive panel fields. */
void orc_gradient(const orc_t* o, const double* a, double* gx, double* ge, double* gr) {
  const int N = o->N, R = o->R;
  const size_t nf = (size_t)N * N * R, ns = (size_t)3 * N * R;
  panel_args_t pa[6];
  args_alloc(o, pa);
  for (int p = 0; p < 6; p++) scalar_args(o, a + p * nf, &pa[p]);
  double *xfp = dalloc(ns), *xfn = dalloc(ns), *yfp = dalloc(ns), *yfn = dalloc(ns);
  double *dx = dalloc(nf), *de = dalloc(nf), *dr = dalloc(nf);
  for (int p = 0; p < 6; p++) {
    exchange(o, pa, p, xfp, xfn, yfp, yfn);
    ddx(o, a + p * nf, dx, xfp, xfn);
    ddy(o, a + p * nf, de, yfp, yfn);
    ddr_cols(o, a + p * nf, dr, (size_t)N * N);
    double *ox = gx + p * nf, *oe = ge + p * nf, *orr = gr + p * nf;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < R; k++)
      for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
          const size_t q = IDX(N, i, j, k);
          const int ij = j * N + i;
          ox[q] = (1.0 / o->gr[k]) * (o->mD[j] * dx[q] + (o->XY[ij] * de[q]) / o->mD[j]);
          oe[q] = (1.0 / o->gr[k]) * (o->mC[i] * de[q] + (o->XY[ij] * dx[q]) / o->mC[i]);
          orr[q] = dr[q];
        }
  }
  free(xfp); free(xfn); free(yfp); free(yfn); free(dx); free(de); free(dr);
  args_free(pa);
}

/* grad.f90:62-136 on all six panels. */
void orc_divergence(const orc_t* o, const double* axi, const double* aeta, const double* ar, double* div) {
  const int N = o->N, R = o->R;
  const size_t nf = (size_t)N * N * R, ns = (size_t)3 * N * R;
  panel_args_t pa[6];
  args_alloc(o, pa);
  for (int p = 0; p < 6; p++) vector_args(o, p + 1, axi + p * nf, aeta + p * nf, &pa[p]);
  double *xfp = dalloc(ns), *xfn = dalloc(ns), *yfp = dalloc(ns), *yfn = dalloc(ns);
  double *dx = dalloc(nf), *de = dalloc(nf), *dr = dalloc(nf);
  for (int p = 0; p < 6; p++) {
    exchange(o, pa, p, xfp, xfn, yfp, yfn);
    ddx(o, axi + p * nf, dx, xfp, xfn);
    ddy(o, aeta + p * nf, de, yfp, yfn);
    ddr_cols(o, ar + p * nf, dr, (size_t)N * N);
    const double *vx = axi + p * nf, *ve = aeta + p * nf, *vr = ar + p * nf;
    double* dv = div + p * nf;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < R; k++)
      for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
          const size_t q = IDX(N, i, j, k);
          const int ij = j * N + i;
          dv[q] = (1.0 / o->gr[k]) * (((o->mdelta[ij] / o->CD[ij]) * (dx[q] / o->mC[i] + de[q] / o->mD[j]) -
                                       (tan(o->gxi[i]) / o->mD[j]) * vx[q]) -
                                      (tan(o->geta[j]) / o->mC[i]) * ve[q]) +
                  dr[q] + (2.0 * vr[q]) / o->gr[k];
        }
  }
  free(xfp); free(xfn); free(yfp); free(yfn); free(dx); free(de); free(dr);
  args_free(pa);
}

/* --------------------------------------------------------------- RHS */

/* compute_rhs_acoustic.f90:18-149 without the forcing block (:92-125).
 * q, rhs: 6 panels x 5 fields, layout q(N,N,R,5,6). */
void orc_compute_rhs(const orc_t* o, const double* q, double* rhs) {
  const int N = o->N, R = o->R;
  const size_t nf = (size_t)N * N * R, nq = 5 * nf;
  double *vxi = dalloc(6 * nf), *veta = dalloc(6 * nf), *vr = dalloc(6 * nf), *p = dalloc(6 * nf);
  for (int z = 0; z < 6; z++) {
    memcpy(vxi + z * nf, q + z * nq + 1 * nf, nf * sizeof(double));
    memcpy(veta + z * nf, q + z * nq + 2 * nf, nf * sizeof(double));
    memcpy(vr + z * nf, q + z * nq + 3 * nf, nf * sizeof(double));
    memcpy(p + z * nf, q + z * nq + 4 * nf, nf * sizeof(double));
  }
  double *div = dalloc(6 * nf), *gx = dalloc(6 * nf), *ge = dalloc(6 * nf), *gr = dalloc(6 * nf);
  orc_divergence(o, vxi, veta, vr, div);
  orc_gradient(o, p, gx, ge, gr);
  /* DDR(rho0, drho0dr) (:51): rho0 is radial-only, one column suffices. */
  double* drho0dr = dalloc(R);
  ddr_cols(o, o->rho0, drho0dr, 1);
  for (int z = 0; z < 6; z++) {
    const double* qz = q + z * nq;
    double* rz = rhs + z * nq;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < R; k++)
      for (size_t c = 0; c < (size_t)N * N; c++) {
        const size_t x = (size_t)k * N * N + c;
        const double rho0 = o->rho0[k], g0 = o->g0[k], c2 = o->c2f ? o->c2f[z * nf + x] : o->c2[k];
        const double dv = div[z * nf + x];
        double r1 = -(rho0 * dv);
        r1 = r1 - qz[3 * nf + x] * drho0dr[k];
        double r2 = -(gx[z * nf + x] / rho0);
        double r3 = -(ge[z * nf + x] / rho0);
        double r4 = -(gr[z * nf + x] / rho0);
        r4 = r4 - (qz[x] * g0) / rho0;
        double r5 = -((c2 * rho0) * dv);
        r5 = r5 + (rho0 * g0) * qz[3 * nf + x];
        rz[x] = r1;
        rz[nf + x] = r2;
        rz[2 * nf + x] = r3;
        rz[3 * nf + x] = r4;
        rz[4 * nf + x] = r5;
      }
  }
  free(vxi); free(veta); free(vr); free(p); free(div); free(gx); free(ge); free(gr); free(drho0dr);
}

/* bc_acoustic_wave.f90:14-34 (rrank=0=pr-1). */
void orc_bc_rhs(const orc_t* o, double* rhs) {
  const int N = o->N, R = o->R;
  if (R == 1) return; /* bc_acoustic_wave.f90:11 */
  const size_t nf = (size_t)N * N * R;
  for (int z = 0; z < 6; z++)
    for (int l = 0; l < 5; l++) {
      double* f = rhs + (size_t)(z * 5 + l) * nf;
      memset(f + (size_t)(R - 1) * N * N, 0, sizeof(double) * N * N);
      memset(f, 0, sizeof(double) * N * N);
    }
}

/* sponge.f90:20-50: single-precision ratio, applied to fields 1, 4, 5. */
double orc_sponge_sigma(const orc_t* o, int gk /* 1-based */) {
  const int L = o->sponge_len;
  if (o->R == 1) return 0.0; /* sponge.f90:18: no sponge in 2-D */
  const int lower = L, upper = o->R - L + 1;
  volatile float num, den;
  if (gk >= upper) {
    num = (float)(gk - upper);
    den = (float)L;
    float ratio = num / den;
    return o->a_sponge * pow((double)ratio, 3.0);
  } else if (gk <= lower) {
    num = (float)(lower - gk);
    den = (float)L;
    float ratio = num / den;
    return o->a_sponge * pow((double)ratio, 3.0);
  }
  return 0.0;
}

void orc_sponge(const orc_t* o, const double* q, double* rhs) {
  const int N = o->N, R = o->R;
  if (R == 1) return; /* sponge.f90:18 */
  const size_t nf = (size_t)N * N * R;
  for (int k = 0; k < R; k++) {
    const double sg = orc_sponge_sigma(o, k + 1);
    for (int z = 0; z < 6; z++) {
      const int fl[3] = {0, 3, 4};
      for (int a = 0; a < 3; a++) {
        const size_t base = (size_t)(z * 5 + fl[a]) * nf + (size_t)k * N * N;
        for (size_t c = 0; c < (size_t)N * N; c++) rhs[base + c] = rhs[base + c] - sg * (q[base + c] - 0.0);
      }
    }
  }
}

/* --------------------------------------------------------------- filter */

/* FILTER_XI core (filter.f90:176-315) for one panel: ghosts gp, gn (3,N,R). */
static void filter_xi(const orc_t* o, const double* g, double* gf, const double* gp, const double* gn) {
  const int N = o->N, R = o->R;
#define GP(l, j, k) gp[((size_t)(k) * N + (j)) * 3 + (l)-1]
#define GN(l, j, k) gn[((size_t)(k) * N + (j)) * 3 + (l)-1]
#pragma omp parallel for schedule(static)
  for (int k = 0; k < R; k++)
    for (int j = 0; j < N; j++) {
      const double* gl = g + IDX(N, 0, j, k);
      double* ol = gf + IDX(N, 0, j, k);
      for (int i = NST + 1; i <= N - NST; i++) {
        double s = 0.0;
        for (int it = -NST; it <= NST; it++) s = s + filt_coeffs[it + 3] * gl[i + it - 1];
        ol[i - 1] = s;
      }
      for (int i = 1; i <= NST; i++) {
        double s = 0.0;
        for (int it = -NST; it <= -i; it++) s = s + filt_coeffs[it + 3] * GP(-i - it + 1, j, k);
        for (int it = -i + 1; it <= NST; it++) s = s + filt_coeffs[it + 3] * gl[i + it - 1];
        ol[i - 1] = s;
      }
      for (int i = N - NST + 1; i <= N; i++) {
        double s = 0.0;
        for (int it = -NST; it <= N - i; it++) s = s + filt_coeffs[it + 3] * gl[i + it - 1];
        for (int it = N - i + 1; it <= NST; it++) s = s + filt_coeffs[it + 3] * GN(i + it - N, j, k);
        ol[i - 1] = s;
      }
    }
#undef GP
#undef GN
}

/* FILTER_ETA core (filter.f90:448-629): ghosts gp, gn (N,3,R). */
static void filter_eta(const orc_t* o, const double* g, double* gf, const double* gp, const double* gn) {
  const int N = o->N, R = o->R;
#define GP(i, l, k) gp[((size_t)(k) * 3 + (l)-1) * N + (i)]
#define GN(i, l, k) gn[((size_t)(k) * 3 + (l)-1) * N + (i)]
#pragma omp parallel for schedule(static)
  for (int k = 0; k < R; k++)
    for (int i = 0; i < N; i++) {
      for (int j = NST + 1; j <= N - NST; j++) {
        double s = 0.0;
        for (int it = -NST; it <= NST; it++) s = s + filt_coeffs[it + 3] * g[IDX(N, i, j + it - 1, k)];
        gf[IDX(N, i, j - 1, k)] = s;
      }
      for (int j = 1; j <= NST; j++) {
        double s = 0.0;
        for (int it = -NST; it <= -j; it++) s = s + filt_coeffs[it + 3] * GP(i, -it - j + 1, k);
        for (int it = -j + 1; it <= NST; it++) s = s + filt_coeffs[it + 3] * g[IDX(N, i, it + j - 1, k)];
        gf[IDX(N, i, j - 1, k)] = s;
      }
      for (int j = N - NST + 1; j <= N; j++) {
        double s = 0.0;
        for (int it = -NST; it <= N - j; it++) s = s + filt_coeffs[it + 3] * g[IDX(N, i, it + j - 1, k)];
        for (int it = N - j + 1; it <= NST; it++) s = s + filt_coeffs[it + 3] * GN(i, it + j - N, k);
        gf[IDX(N, i, j - 1, k)] = s;
      }
    }
#undef GP
#undef GN
}

/* FILTER_R (filter.f90:633-781), pr=1: the closure loops are commented out
 * in the reference, so the 5 planes at each end stay 0. */
static void filter_r(const orc_t* o, const double* g, double* gf) {
  const int N = o->N, R = o->R;
  const size_t plane = (size_t)N * N;
  if (R == 1) { /* 2-D case, filter.f90:647-650: gf = g */
    memcpy(gf, g, sizeof(double) * plane);
    return;
  }
  memset(gf, 0, sizeof(double) * plane * R);
#pragma omp parallel for schedule(static)
  for (int k = NSTR + 1; k <= R - NSTR; k++)
    for (size_t c = 0; c < plane; c++) {
      double s = 0.0;
      for (int it = -NSTR; it <= NSTR; it++) s = s + filt_coeffs_r[it + 5] * g[(size_t)(k + it - 1) * plane + c];
      gf[(size_t)(k - 1) * plane + c] = s;
    }
}

/* filter.f90:9-48 FILTER_EXP on q(N,N,R,5,6) -> qf. */
void orc_filter_exp(const orc_t* o, const double* q, double* qf) {
  const int N = o->N, R = o->R;
  const size_t nf = (size_t)N * N * R, ns = (size_t)3 * N * R;
  double *t1 = dalloc(6 * nf), *t2 = dalloc(6 * nf);
  double *xfp = dalloc(ns), *xfn = dalloc(ns), *yfp = dalloc(ns), *yfn = dalloc(ns);
  panel_args_t pa[6];
  args_alloc(o, pa);
  for (int m = 0; m < 5; m++) {
    for (int z = 0; z < 6; z++) scalar_args(o, q + (size_t)(z * 5 + m) * nf, &pa[z]);
    for (int z = 0; z < 6; z++) {
      exchange(o, pa, z, xfp, xfn, yfp, yfn);
      filter_xi(o, q + (size_t)(z * 5 + m) * nf, t1 + z * nf, xfp, xfn);
      filter_eta(o, t1 + z * nf, t2 + z * nf, yfp, yfn);
      filter_r(o, t2 + z * nf, qf + (size_t)(z * 5 + m) * nf);
    }
  }
  args_free(pa);
  free(t1); free(t2); free(xfp); free(xfn); free(yfp); free(yfn);
}

/* ---------------------------------------------------------------- RK4 */

/* RHS as one stage sees it: COMPUTE_RHS, BC_RHS, SPONGE (timestep.f90:64-66). */
void orc_stage_rhs(const orc_t* o, const double* q, double* rhs) {
  orc_compute_rhs(o, q, rhs);
  orc_bc_rhs(o, rhs);
  orc_sponge(o, q, rhs);
}

/* timestep.f90:41-86 RK4, repeated nsteps times. q updated in place. */
void orc_rk4(const orc_t* o, double* q, int nsteps, double dt, double* time) {
  const size_t nq = (size_t)30 * o->N * o->N * o->R;
  double *rhs = dalloc(nq), *qw = dalloc(nq), *qf = dalloc(nq);
  for (int step = 0; step < nsteps; step++) {
    memset(qw, 0, nq * sizeof(double));
    for (int s = 0; s < 5; s++) {
      orc_stage_rhs(o, q, rhs);
#pragma omp parallel for schedule(static)
      for (size_t x = 0; x < nq; x++) {
        qw[x] = alpha_rk4[s] * qw[x] + dt * rhs[x];
        q[x] = q[x] + beta_rk4[s] * qw[x];
      }
    }
    orc_filter_exp(o, q, qf);
    memcpy(q, qf, nq * sizeof(double));
    if (time) *time = *time + dt;
  }
  free(rhs); free(qw); free(qf);
}

/* ------------------------------------------------------ test entry points */

/* Ghost values every panel receives for a scalar field f (6 panel fields):
 * gx (6,2,R,N,3) [fp,fn of DDX], gy (6,2,R,3,N) [fp,fn of DDY]. */
void orc_scalar_ghosts(const orc_t* o, const double* f, double* gx, double* gy) {
  const size_t nf = (size_t)o->N * o->N * o->R, ns = (size_t)3 * o->N * o->R;
  panel_args_t pa[6];
  args_alloc(o, pa);
  for (int p = 0; p < 6; p++) scalar_args(o, f + p * nf, &pa[p]);
  for (int p = 0; p < 6; p++)
    exchange(o, pa, p, gx + (2 * p) * ns, gx + (2 * p + 1) * ns, gy + (2 * p) * ns, gy + (2 * p + 1) * ns);
  args_free(pa);
}

/* Receiver-basis ghosts of the divergence for a vector field (vxi, veta):
 * gx = v_xi ghosts of DDX, gy = v_eta ghosts of DDY, shapes as above. */
void orc_vector_ghosts(const orc_t* o, const double* vxi, const double* veta, double* gx, double* gy) {
  const size_t nf = (size_t)o->N * o->N * o->R, ns = (size_t)3 * o->N * o->R;
  panel_args_t pa[6];
  args_alloc(o, pa);
  for (int p = 0; p < 6; p++) vector_args(o, p + 1, vxi + p * nf, veta + p * nf, &pa[p]);
  for (int p = 0; p < 6; p++)
    exchange(o, pa, p, gx + (2 * p) * ns, gx + (2 * p + 1) * ns, gy + (2 * p) * ns, gy + (2 * p + 1) * ns);
  args_free(pa);
}

/* Single-panel operators with caller-supplied ghosts, for closed-form tests. */
void orc_ddx(const orc_t* o, const double* f, double* d, const double* fp, const double* fn) { ddx(o, f, d, fp, fn); }
void orc_ddy(const orc_t* o, const double* f, double* d, const double* fp, const double* fn) { ddy(o, f, d, fp, fn); }
void orc_ddr(const orc_t* o, const double* f, double* d, long ncol) { ddr_cols(o, f, d, (size_t)ncol); }
void orc_filter_xi(const orc_t* o, const double* g, double* gf, const double* gp, const double* gn) { filter_xi(o, g, gf, gp, gn); }
void orc_filter_eta(const orc_t* o, const double* g, double* gf, const double* gp, const double* gn) { filter_eta(o, g, gf, gp, gn); }
void orc_filter_r(const orc_t* o, const double* g, double* gf) { filter_r(o, g, gf); }
void orc_interp_x(const orc_t* o, const double* in, double* out) { interp_x(o, in, out); }
void orc_interp_y(const orc_t* o, const double* in, double* out) { interp_y(o, in, out); }
void orc_convert(const orc_t* o, int z, const double* axi, const double* aeta, double* ath, double* aph) {
  convert_to_theta_phi(o, z, axi, aeta, ath, aph);
}
void orc_rk_coeffs(double* a, double* b) {
  for (int s = 0; s < 5; s++) { a[s] = alpha_rk4[s]; b[s] = beta_rk4[s]; }
}
