// Interface between the host C-ABI layer and the device kernels.
#pragma once
#include <hip/hip_runtime_api.h>

#include "solar_tables.hpp"

namespace solar {

// Device-resident views of the tables (all pointers are device pointers).
struct DevTables {
  int n, nr;
  const double* ci;      // metric_C(i)            [n]
  const double* dj;      // metric_D(j)            [n]
  const double* dc;      // metric_delta/CD        [n*n]
  const double* tdx;     // tan(xi)/metric_D       [n*n]
  const double* tde;     // tan(eta)/metric_C      [n*n]
  const double* xy;      // XY                     [n*n]
  const RadialRow* rad;  // per-plane constants    [nr]
  const double* conv;    // basis conversion       [3][n*n][6]
  const double* gw;      // gather weights         [3][n][9]
  const double* trans;   // receiver transform     [6][4][3][n][2]
  const Route* route;    // panel routes           [6][4]
  const int* gt_mp;      // transposed gather rows [3][n][kGtWidth]: receiver along index
  const double* gt_w;    //   and weight per entry
  const double* adj_metric;  // adjoint metric factors [n*n][4]
  int inv_route[24];     // (receiver P*4+e) | reverse << 8 per (sender, sender edge)
  double fcx[7], fcy[7];
  double filt[7], filtr[11];
  int fast_div;  // every divisor of the stage (C, D, r, rho0) lies in [2^-60, 2^60]
};

// One member of a batched ensemble (FP64): its tables, state ping-pong
// buffers, 2N accumulator, ghost strips, edge-column copies (written by
// every stage, read by the next gather, edge_cols_elems(n, nr) values) and
// time step.
struct EnsMember {
  DevTables T;
  double* q[2];
  double* qw;
  double *gxp, *gyp, *gxv, *gyv;
  double *fgx, *fgy;
  double* ecols;
  double dt;
};

// Launch-shape overrides (0 = automatic).
struct LaunchShape {
  int stage_chunks = 0;
  int filter_chunks = 0;
  int stage_ti = 0, stage_tj = 0;
  int filter_ti = 0, filter_tj = 0;
  int adj_stage_chunks = 0, adj_filter_chunks = 0;  // sensitivity kernels
};

// State pointers are float* when f32 is set (FP32-state variant), double* otherwise.
// ecols (optional): the edge_cols copies of q's W/E edge columns; the
// column strips are then read from there.
// qwr (optional, FP64 storage): the gathered field is q + betar * qwr, a
// stage input rebuilt from q and qw (kStageRecon); ecols then holds it as
// computed.
hipError_t launch_gather(const DevTables& T, const void* q, bool f32, bool vec, int sfield, int nfields, double* gxs,
                         double* gys, double* gxv, double* gyv, hipStream_t st, const void* ecols = nullptr,
                         const void* qwr = nullptr, double betar = 0.0);
// c2f (optional): a 3-D c^2 field (n,n,nr,6) used instead of the radial
// profile; divout (optional): receives div(v) of the stage input (n,n,nr,6),
// as double, or as float when div32 is set.  tl_dc2, tl_div (optional, both
// or neither): tangent-linear stage, adding -dc2*rho0*div_primal to the
// pressure equation (dc2 and div_primal fields (n,n,nr,6)).  ecols
// (optional, not with rhs_mode): receives the output's edge columns in the
// edge_cols layout (p', v_xi, v_eta; all five fields after the last stage).
// xf (FP64 storage, not with rhs_mode or tl_dc2): kStoreNoQ, the stage
// stores qw in place and the edge columns of its output, not q (qout unused);
// kStageRecon (stages 2-4), the stage input is qin + beta_r * qw, its qw goes
// to qwo (not qw) and its q to qout.  Tile shapes other than 32x8 and 16x16
// are not compiled for these: requested ones fall back to the default.
constexpr int kStoreNoQ = 1, kStageRecon = 2;
hipError_t launch_stage(const DevTables& T, const void* qin, void* qout, void* qw, bool f32, const double* gxp,
                        const double* gyp, const double* gxv, const double* gyv, int stage, double dt, int rhs_mode,
                        hipStream_t st, const LaunchShape& ls, const double* c2f = nullptr,
                        double* divout = nullptr, int div32 = 0, const double* tl_dc2 = nullptr,
                        const double* tl_div = nullptr, void* ecols = nullptr, int xf = 0, void* qwo = nullptr,
                        double beta_r = 0.0);
hipError_t launch_filter(const DevTables& T, const void* qin, void* qout, bool f32, const double* fgx,
                         const double* fgy, hipStream_t st, const LaunchShape& ls);
// Ensemble launches.  items: 4 ints per work item, (member, bi | bj << 16,
// P (stage) or 5P+F (filter) | k0 << 8, k1); grid blocks walk the list with a
// grid-sized stride.  Stage ti x tj in {64x8, 32x8, 16x16}, filter in
// {32x16, 16x16, 32x8}.  The state goes from M.q[cur] to M.q[1-cur].
hipError_t launch_stage_ens(const EnsMember* M, const int* items, int nitems, int grid, int ti, int tj, int stage,
                            int cur, hipStream_t st);
hipError_t launch_filter_ens(const EnsMember* M, const int* items, int nitems, int grid, int ti, int tj, int cur,
                             hipStream_t st);
// A gather over a group of members: table[x] = first block of the group's
// member x (24 * gather_blocks_per_edge(nr, chunks) blocks each), table[nmem +
// x] = its index in M; nblocks in total; maxn: the largest n of the group
// (block size and LDS).  vec:
// the stage ghosts (p', velocity), else the filter's five scalars.  cols:
// the W/E column strips come from the members' edge-column copies (every
// gather but the first of a step, whose source no stage wrote).
hipError_t launch_gather_ens(const EnsMember* M, const int* table, int nmem, int nblocks, int maxn, bool vec, int cur,
                             bool cols, int chunks, hipStream_t st);
// Planes per gather block (about 32 chunks per receiver edge), and gather
// blocks per receiver edge.
constexpr int gather_kc(int nr, int chunks = 32) { return (nr + chunks - 1) / chunks; }
constexpr int gather_blocks_per_edge(int nr, int chunks = 32) {
  return (nr + gather_kc(nr, chunks) - 1) / gather_kc(nr, chunks);
}
// Threads per gather block: one strip point each up to 1024 (n <= 341), else
// two each (n <= 512).
constexpr int gather_threads_max(int ni) { return ni == 1 ? 1024 : 768; }
inline int gather_threads(int n, int ni) { return (((3 * n + ni - 1) / ni) + 63) / 64 * 64; }
// Compact copies of the edge columns a gather reads from the W/E edges
// (columns i = 1..3 and n-4..n-2 of every row): edge_cols[P][side W/E][k]
// [field 0..4][j][layer 1..3], so one strip of one field and plane is 3n
// contiguous values.  Elements per (panel, side, plane), and in all.
__host__ __device__ constexpr size_t edge_cols_plane(int n) { return (size_t)15 * n; }
inline size_t edge_cols_elems(int n, int nr) { return (size_t)12 * nr * edge_cols_plane(n); }
// mode 0: count of non-finite values -> acc[0]; mode 1: sum of squares of
// `field` over all panels -> acc[1].
hipError_t launch_scan(const void* q, bool f32, size_t nf, double* acc, int mode, hipStream_t st, int field = 3);
hipError_t launch_place(const double* src, void* dst, bool f32, size_t nf, int dfield, hipStream_t st);

}  // namespace solar
