// C ABI of the acoustic RK4 hot path (declarations and reference citations in
// include/solar_amd.h).  One handle owns the device tables, the state
// q(n,n,nr,5,6) and its work buffers on one HIP device and one stream.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/solar_amd.h"
#include "solar_adjoint.hpp"
#include "solar_internal.hpp"
#include "solar_kernels.hpp"
#include "solar_tables.hpp"

using solar::DevTables;
using solar::LaunchShape;
using solar::RadialRow;
using solar::Route;
using solar::Tables;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return fail(SOLAR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

struct solar_ctx {
  int n = 0, nr = 0, device = 0;
  size_t nf = 0;  // points per field per panel
  bool f32 = false;  // state stored as float (FP32-state variant)
  size_t esz = 8;    // bytes per state element
  hipStream_t stream = nullptr;
  Tables tab;
  DevTables dt{};
  std::vector<void*> allocs;
  void* q[2] = {nullptr, nullptr};  // state ping-pong, element type per f32
  void* qw = nullptr;               // 2N accumulator, same type
  void* qw2 = nullptr;              // second accumulator of the stored-q schedule (FP64 state)
  double *gxp = nullptr, *gyp = nullptr, *gxv = nullptr, *gyv = nullptr;
  double *fgx = nullptr, *fgy = nullptr;
  void* ecols = nullptr;  // edge-column copies written by the stages for the next gather (state type)
  double* acc = nullptr;
  LaunchShape shape;
  // hipGraph of one RK4 step for a given dt
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  double graph_dt = 0.0;
  // event timing
  bool profile = false;
  std::vector<hipEvent_t> ev;
  double ktime[5] = {0, 0, 0, 0, 0};
  size_t bytes = 0;
  // sensitivity / adjoint state, allocated on first use
  double* c2f = nullptr;  // optional 3-D c^2 (n,n,nr,6); points at c2f_buf when set
  double* c2f_buf = nullptr;
  double* c2f_graph = nullptr;
  double *wb[2] = {nullptr, nullptr}, *tmp = nullptr, *lam = nullptr, *gc2 = nullptr;
  long long* rec = nullptr;  // receiver indices into the state (field p')
  int div_bits = 64;         // storage of the stage divergences kept for the gradient
  bool div_cache = true;     // keep trailing segments' divergences from the forward pass in free memory
  void* adj_ws = nullptr;    // workspace of solar_misfit_gradient (checkpoints, divergences, traces), grow-only
  size_t adj_ws_bytes = 0;
  double adj_times[4] = {0, 0, 0, 0};  // last gradient: workspace, forward, recompute, adjoint (s)
  int adj_recomputed = 0;              // last gradient: steps recomputed from checkpoints
  int nrec = 0, rec_cap = 0;
  // static device scratch of solar_compute_rhs_noalloc on this handle's device
  void* scratch = nullptr;
  hipError_t scratch_status = hipSuccess;
};

namespace {

template <typename T>
int dev_alloc(solar_ctx* c, T** p, size_t count) {
  void* v = nullptr;
  const size_t b = count * sizeof(T);
  hipError_t e = hipMalloc(&v, b);
  if (e != hipSuccess)
    return fail(SOLAR_ERR_HIP, "hipMalloc(" + std::to_string(b) + " bytes): " + hipGetErrorString(e));
  c->allocs.push_back(v);
  c->bytes += b;
  *p = static_cast<T*>(v);
  return SOLAR_OK;
}

// Frees one buffer dev_alloc made (bytes long) and drops it from the handle's
// list and byte count.  The caller has synchronised the handle's stream.
void dev_free(solar_ctx* c, void* p, size_t bytes) {
  if (!p) return;
  for (size_t a = 0; a < c->allocs.size(); ++a)
    if (c->allocs[a] == p) {
      c->allocs.erase(c->allocs.begin() + static_cast<long>(a));
      c->bytes -= bytes;
      (void)hipFree(p);
      return;
    }
}

void release(solar_ctx* c) {
  if (c->adj_ws) (void)hipFree(c->adj_ws);
  c->adj_ws = nullptr;
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  c->allocs.clear();
}

// Enqueues one RK4 step (timestep.f90:61-82) on the handle's stream.  The
// state lives in q[0] before and after: five stages ping-pong q[0] <-> q[1],
// the filter brings it back.  With `timed`, events bracket every launch.
// divp: where stage s stores div(v) of its input.  b0, b1, bw: other state
// and accumulator buffers (a tangent-linear state); tl_dc2 with tl_divp: the
// tangent-linear source of each stage (primal divergences tl_divp[s]).
//
// With an FP64 state the step stores q after stages 2, 4 and 5 only (720
// instead of 800 bytes per point): stages 1 and 3 store their accumulator
// and the edge columns of q, and stages 2 and 4 rebuild their input as
// q + beta*qw wherever they read it, with the expression that produced it,
// so every value is the same as when q is stored.  Buffers, A = q[0], B =
// q[1], X = qw, Y = qw2 (a rebuilding stage reads its qw at other blocks'
// halo cells, so it writes its own to the other accumulator):
//   stage 1  A -> qw X                 stage 2  A + b1 X -> q B, qw Y
//   stage 3  B, Y -> qw Y              stage 4  B + b3 Y -> q A, qw X
//   stage 5  A, X -> q B               filter   B -> A
int enqueue_step(solar_ctx* c, double dt, bool timed, double* const* divp = nullptr, void* b0 = nullptr,
                 void* b1 = nullptr, void* bw = nullptr, const double* tl_dc2 = nullptr,
                 double* const* tl_divp = nullptr) {
  const DevTables& T = c->dt;
  void* cur = b0 ? b0 : c->q[0];
  void* nxt = b1 ? b1 : c->q[1];
  void* qw = bw ? bw : c->qw;
  int e = 0;
  if (c->qw2 && !b0 && !tl_dc2) {
    void* A = c->q[0];
    void* B = c->q[1];
    void* X = c->qw;
    void* Y = c->qw2;
    // per stage: gathered input (q, rebuilt from qw with beta), stage buffers
    struct St {
      void *qg, *wg;
      double bg;
      void *qin, *qout, *qw, *qwo;
      int xf;
    };
    const St sch[5] = {
        {A, nullptr, 0.0, A, B, X, nullptr, solar::kStoreNoQ},  // (q out: not stored)
        {A, X, solar::kBetaRk4[0], A, B, X, Y, solar::kStageRecon},
        {B, nullptr, 0.0, B, A, Y, nullptr, solar::kStoreNoQ},
        {B, Y, solar::kBetaRk4[2], B, A, Y, X, solar::kStageRecon},
        {A, nullptr, 0.0, A, B, X, nullptr, 0},
    };
    for (int s = 0; s < 5; ++s) {
      const St& x = sch[s];
      if (timed) HIPCHK(hipEventRecord(c->ev[e++], c->stream));
      HIPCHK(solar::launch_gather(T, x.qg, false, true, 4, 1, c->gxp, c->gyp, c->gxv, c->gyv, c->stream,
                                  s > 0 ? c->ecols : nullptr, x.wg, x.bg));
      if (timed) HIPCHK(hipEventRecord(c->ev[e++], c->stream));
      HIPCHK(solar::launch_stage(T, x.qin, x.qout, x.qw, false, c->gxp, c->gyp, c->gxv, c->gyv, s, dt,
                                 0, c->stream, c->shape, c->c2f, divp ? divp[s] : nullptr, c->div_bits == 32,
                                 nullptr, nullptr, c->ecols, x.xf, x.qwo, x.bg));
    }
    if (timed) HIPCHK(hipEventRecord(c->ev[e++], c->stream));
    HIPCHK(solar::launch_gather(T, B, false, false, 0, 5, c->fgx, c->fgy, nullptr, nullptr, c->stream, c->ecols));
    if (timed) HIPCHK(hipEventRecord(c->ev[e++], c->stream));
    HIPCHK(solar::launch_filter(T, B, A, false, c->fgx, c->fgy, c->stream, c->shape));
    if (timed) HIPCHK(hipEventRecord(c->ev[e++], c->stream));
    return SOLAR_OK;
  }
  // Every stage also writes its output's edge columns to c->ecols, from which
  // the next gather reads its W/E strips; the first gather reads the state.
  for (int s = 0; s < 5; ++s) {
    if (timed) HIPCHK(hipEventRecord(c->ev[e++], c->stream));
    HIPCHK(solar::launch_gather(T, cur, c->f32, true, 4, 1, c->gxp, c->gyp, c->gxv, c->gyv, c->stream,
                                s > 0 ? c->ecols : nullptr));
    if (timed) HIPCHK(hipEventRecord(c->ev[e++], c->stream));
    HIPCHK(solar::launch_stage(T, cur, nxt, qw, c->f32, c->gxp, c->gyp, c->gxv, c->gyv, s, dt, 0, c->stream,
                               c->shape, c->c2f, divp ? divp[s] : nullptr, c->div_bits == 32, tl_dc2,
                               tl_dc2 ? tl_divp[s] : nullptr, c->ecols));
    void* t = cur;
    cur = nxt;
    nxt = t;
  }
  if (timed) HIPCHK(hipEventRecord(c->ev[e++], c->stream));
  HIPCHK(solar::launch_gather(T, cur, c->f32, false, 0, 5, c->fgx, c->fgy, nullptr, nullptr, c->stream, c->ecols));
  if (timed) HIPCHK(hipEventRecord(c->ev[e++], c->stream));
  HIPCHK(solar::launch_filter(T, cur, nxt, c->f32, c->fgx, c->fgy, c->stream, c->shape));
  if (timed) HIPCHK(hipEventRecord(c->ev[e++], c->stream));
  return SOLAR_OK;
}

int build_graph(solar_ctx* c, double dt) {
  if (c->graph_exec) {
    (void)hipGraphExecDestroy(c->graph_exec);
    c->graph_exec = nullptr;
  }
  if (c->graph) {
    (void)hipGraphDestroy(c->graph);
    c->graph = nullptr;
  }
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_step(c, dt, false);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) return rc;
  if (e != hipSuccess) return fail(SOLAR_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  c->graph = g;
  HIPCHK(hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0));
  c->graph_dt = dt;
  c->c2f_graph = c->c2f;
  return SOLAR_OK;
}

bool env_flag(const char* name) {
  const char* v = std::getenv(name);
  return v && v[0] && std::strcmp(v, "0") != 0;
}

int env_int(const char* name) {
  const char* v = std::getenv(name);
  return v ? std::atoi(v) : 0;
}

}  // namespace

int solar::set_error(int code, const std::string& msg) { return fail(code, msg); }
void solar::clear_error() { g_err.clear(); }

int solar::upload_tables(const Tables& t, DevTables* d, std::vector<void*>* allocs, size_t* bytes) {
  auto put = [&](auto** dst, const auto& src) -> int {
    using T = std::remove_cv_t<std::remove_reference_t<decltype(src[0])>>;
    void* v = nullptr;
    const size_t b = src.size() * sizeof(T);
    hipError_t e = hipMalloc(&v, b);
    if (e != hipSuccess) return fail(SOLAR_ERR_HIP, "hipMalloc(" + std::to_string(b) + " bytes): " + hipGetErrorString(e));
    allocs->push_back(v);
    *bytes += b;
    e = hipMemcpy(v, src.data(), b, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(SOLAR_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    *dst = static_cast<T*>(v);
    return SOLAR_OK;
  };
  int rc;
  d->n = t.n;
  d->nr = t.nr;
  // the stage divides by C, D, r and rho0 through their reciprocals when all
  // lie in [2^-60, 2^60] (quotients of normal numerators then stay normal)
  auto in_range = [](double v) { return v >= 0x1p-60 && v <= 0x1p60; };
  bool fast = true;
  for (double v : t.metric_c) fast = fast && in_range(v);
  for (double v : t.metric_d) fast = fast && in_range(v);
  for (const auto& row : t.radial) fast = fast && in_range(row.r) && in_range(row.rho0);
  d->fast_div = fast ? 1 : 0;
  if ((rc = put(&d->ci, t.metric_c)) || (rc = put(&d->dj, t.metric_d)) || (rc = put(&d->dc, t.dc)) ||
      (rc = put(&d->tdx, t.tdx)) || (rc = put(&d->tde, t.tde)) || (rc = put(&d->xy, t.xy)) ||
      (rc = put(&d->rad, t.radial)) || (rc = put(&d->conv, t.conv)) || (rc = put(&d->gw, t.gather_w)) ||
      (rc = put(&d->trans, t.trans)))
    return rc;
  std::vector<Route> routes(24);
  for (int p = 0; p < 6; ++p)
    for (int e = 0; e < 4; ++e) routes[p * 4 + e] = t.route[p][e];
  if ((rc = put(&d->route, routes))) return rc;
  if ((rc = put(&d->gt_mp, t.gt_mp)) || (rc = put(&d->gt_w, t.gt_w)) ||
      (rc = put(&d->adj_metric, t.adj_metric)))
    return rc;
  for (int p = 0; p < 6; ++p)
    for (int e = 0; e < 4; ++e) d->inv_route[p * 4 + e] = t.inv_route[p][e];
  for (int q = 0; q < 7; ++q) {
    d->fcx[q] = t.fcx[q];
    d->fcy[q] = t.fcy[q];
    d->filt[q] = solar::kFiltCoeffs[q];
  }
  for (int q = 0; q < 11; ++q) d->filtr[q] = solar::kFiltCoeffsR[q];
  // the stage kernel uses fcx for both angular stencils and skips their zero
  // centre tap
  for (int q = 0; q < 7; ++q)
    if (d->fcx[q] != d->fcy[q]) return fail(SOLAR_ERR_ARG, "xi and eta stencil coefficients differ");
  if (d->fcx[3] != 0.0) return fail(SOLAR_ERR_ARG, "the angular stencil's centre tap is not zero");
  // the filter kernel computes each radial product once for taps q and 10 - q
  for (int q = 0; q < 11; ++q)
    if (d->filtr[q] != d->filtr[10 - q]) return fail(SOLAR_ERR_ARG, "radial filter taps are not symmetric");
  return SOLAR_OK;
}

extern "C" {

const char* solar_last_error(void) { return g_err.c_str(); }



int solar_init(solar_ctx** out, int n, int nr, const double* gr, const double* c2, const double* rho0,
               const double* g0, int sponge_len, double a_sponge) {
  return solar_init_ex(out, n, nr, gr, c2, rho0, g0, sponge_len, a_sponge, 64);
}

int solar_init_ex(solar_ctx** out, int n, int nr, const double* gr, const double* c2, const double* rho0,
                  const double* g0, int sponge_len, double a_sponge, int state_bits) {
  if (!out || !gr || !c2 || !rho0 || !g0) return fail(SOLAR_ERR_ARG, "solar_init: null argument");
  if (state_bits != 64 && state_bits != 32) return fail(SOLAR_ERR_ARG, "solar_init: state_bits must be 64 or 32");
  *out = nullptr;
  solar_ctx* c = new solar_ctx();
  std::string err;
  if (!solar::build_tables(n, nr, gr, c2, rho0, g0, sponge_len, a_sponge, &c->tab, &err)) {
    delete c;
    return fail(SOLAR_ERR_ARG, "solar_init: " + err);
  }
  c->n = n;
  c->nr = nr;
  c->nf = static_cast<size_t>(n) * n * nr;
  c->f32 = state_bits == 32;
  c->esz = c->f32 ? sizeof(float) : sizeof(double);
  int ndev = 0;
  hipError_t he = hipGetDeviceCount(&ndev);
  if (he != hipSuccess || ndev < 1) {
    delete c;
    return fail(SOLAR_ERR_HIP, "solar_init: no HIP device available");
  }
  (void)hipGetDevice(&c->device);
  int rc = SOLAR_OK;
  auto bail = [&](int code) {
    release(c);
    delete c;
    return code;
  };
  if (hipStreamCreate(&c->stream) != hipSuccess) return bail(fail(SOLAR_ERR_HIP, "hipStreamCreate failed"));

  if ((rc = solar::upload_tables(c->tab, &c->dt, &c->allocs, &c->bytes))) return bail(rc);

  const size_t nq = c->nf * 30;
  const size_t ng = static_cast<size_t>(12) * nr * n * 3;  // one field's ghosts, 6 panels x 2 sides
  char *q0 = nullptr, *q1 = nullptr, *qwb = nullptr;
  if ((rc = dev_alloc(c, &q0, nq * c->esz)) || (rc = dev_alloc(c, &q1, nq * c->esz)) ||
      (rc = dev_alloc(c, &qwb, nq * c->esz)) || (rc = dev_alloc(c, &c->gxp, ng)) || (rc = dev_alloc(c, &c->gyp, ng)) || (rc = dev_alloc(c, &c->gxv, ng)) ||
      (rc = dev_alloc(c, &c->gyv, ng)) || (rc = dev_alloc(c, &c->fgx, 5 * ng)) ||
      (rc = dev_alloc(c, &c->fgy, 5 * ng)) || (rc = dev_alloc(c, &c->acc, 2)))
    return bail(rc);
  c->q[0] = q0;
  c->q[1] = q1;
  c->qw = qwb;
  if (!c->f32 && !env_flag("SOLAR_STEP_STORE_ALL")) {  // the stored-q schedule's second accumulator
    char* qw2 = nullptr;
    if ((rc = dev_alloc(c, &qw2, nq * c->esz))) return bail(rc);
    c->qw2 = qw2;
  }
  if (!env_flag("SOLAR_NO_EDGE_COLS")) {
    char* ec = nullptr;
    if ((rc = dev_alloc(c, &ec, solar::edge_cols_elems(n, nr) * c->esz))) return bail(rc);
    c->ecols = ec;
  }
  if (hipMemset(c->q[0], 0, nq * c->esz) != hipSuccess) return bail(fail(SOLAR_ERR_HIP, "hipMemset failed"));
  // address of the no-allocation entry point's device scratch on this device
  // (the current one), looked up once per device
  c->scratch_status = solar::scratch_address(c->device, &c->scratch);
  c->use_graph = !env_flag("SOLAR_NO_GRAPH");
  c->profile = env_flag("SOLAR_PROFILE");
  c->shape.stage_chunks = env_int("SOLAR_STAGE_CHUNKS");
  c->shape.filter_chunks = env_int("SOLAR_FILTER_CHUNKS");
  c->shape.stage_ti = env_int("SOLAR_STAGE_TI");
  c->shape.stage_tj = env_int("SOLAR_STAGE_TJ");
  c->shape.filter_ti = env_int("SOLAR_FILTER_TI");
  c->shape.filter_tj = env_int("SOLAR_FILTER_TJ");
  c->shape.adj_stage_chunks = env_int("SOLAR_ADJ_STAGE_CHUNKS");
  c->shape.adj_filter_chunks = env_int("SOLAR_ADJ_FILTER_CHUNKS");
  if (c->profile) {
    c->ev.resize(13);
    for (auto& e : c->ev)
      if (hipEventCreate(&e) != hipSuccess) return bail(fail(SOLAR_ERR_HIP, "hipEventCreate failed"));
  }
  *out = c;
  g_err.clear();
  return SOLAR_OK;
}

int solar_finalize(solar_ctx* c) {
  if (!c) return SOLAR_OK;
  (void)hipStreamSynchronize(c->stream);
  release(c);
  delete c;
  return SOLAR_OK;
}

// Host double array -> device buffer in the handle's storage type.
static int upload_state(solar_ctx* c, void* dst, const double* src, size_t count) {
  if (!c->f32) {
    HIPCHK(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyHostToDevice, c->stream));
  } else {
    std::vector<float> tmp(src, src + count);
    HIPCHK(hipMemcpyAsync(dst, tmp.data(), count * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return SOLAR_OK;
}

// Device buffer in the handle's storage type -> host double array.
static int download_state(solar_ctx* c, double* dst, const void* src, size_t count) {
  if (!c->f32) {
    HIPCHK(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> tmp(count);
    HIPCHK(hipMemcpyAsync(tmp.data(), src, count * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t x = 0; x < count; ++x) dst[x] = tmp[x];
  }
  return SOLAR_OK;
}

int solar_set_state(solar_ctx* c, const double* q) {
  if (!c || !q) return fail(SOLAR_ERR_ARG, "solar_set_state: null argument");
  int rc = upload_state(c, c->q[0], q, c->nf * 30);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return SOLAR_OK;
}

int solar_get_state(solar_ctx* c, double* q) {
  if (!c || !q) return fail(SOLAR_ERR_ARG, "solar_get_state: null argument");
  return download_state(c, q, c->q[0], c->nf * 30);
}

int solar_set_state_device(solar_ctx* c, const void* q) {
  if (!c || !q) return fail(SOLAR_ERR_ARG, "solar_set_state_device: null argument");
  HIPCHK(hipMemcpyAsync(c->q[0], q, c->nf * 30 * c->esz, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SOLAR_OK;
}

int solar_get_state_device(solar_ctx* c, void* q) {
  if (!c || !q) return fail(SOLAR_ERR_ARG, "solar_get_state_device: null argument");
  HIPCHK(hipMemcpyAsync(q, c->q[0], c->nf * 30 * c->esz, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SOLAR_OK;
}

int solar_state_bits(solar_ctx* c, int* bits) {
  if (!c || !bits) return fail(SOLAR_ERR_ARG, "solar_state_bits: null argument");
  *bits = c->f32 ? 32 : 64;
  return SOLAR_OK;
}

int solar_rk4_async(solar_ctx* c, int nsteps, double dt) {
  if (!c || nsteps < 0) return fail(SOLAR_ERR_ARG, "solar_rk4_async: invalid argument");
  if (c->profile) {
    for (double& v : c->ktime) v = 0.0;
    for (int s = 0; s < nsteps; ++s) {
      int rc = enqueue_step(c, dt, true);
      if (rc) return rc;
      HIPCHK(hipEventSynchronize(c->ev[12]));
      float ms = 0.f;
      for (int st = 0; st < 5; ++st) {
        HIPCHK(hipEventElapsedTime(&ms, c->ev[2 * st], c->ev[2 * st + 1]));
        c->ktime[1] += ms;
        HIPCHK(hipEventElapsedTime(&ms, c->ev[2 * st + 1], c->ev[2 * st + 2]));
        c->ktime[0] += ms;
      }
      HIPCHK(hipEventElapsedTime(&ms, c->ev[10], c->ev[11]));
      c->ktime[3] += ms;
      HIPCHK(hipEventElapsedTime(&ms, c->ev[11], c->ev[12]));
      c->ktime[2] += ms;
      HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[12]));
      c->ktime[4] += ms;
    }
    if (nsteps > 0) {
      c->ktime[0] /= 5.0 * nsteps;
      c->ktime[1] /= 5.0 * nsteps;
      c->ktime[2] /= nsteps;
      c->ktime[3] /= nsteps;
      c->ktime[4] /= nsteps;
    }
    return SOLAR_OK;
  }
  if (c->use_graph) {
    if (!c->graph_exec || c->graph_dt != dt || c->c2f_graph != c->c2f) {
      int rc = build_graph(c, dt);
      if (rc) return rc;
    }
    for (int s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(c->graph_exec, c->stream));
    return SOLAR_OK;
  }
  for (int s = 0; s < nsteps; ++s) {
    int rc = enqueue_step(c, dt, false);
    if (rc) return rc;
  }
  return SOLAR_OK;
}

int solar_sync(solar_ctx* c) {
  if (!c) return fail(SOLAR_ERR_ARG, "solar_sync: null handle");
  HIPCHK(hipStreamSynchronize(c->stream));
  return SOLAR_OK;
}

int solar_rk4(solar_ctx* c, int nsteps, double dt, double* time) {
  int rc = solar_rk4_async(c, nsteps, dt);
  if (rc) return rc;
  rc = solar_sync(c);
  if (rc) return rc;
  if (time)
    for (int s = 0; s < nsteps; ++s) *time = *time + dt;  // timestep.f90:82, once per step
  return SOLAR_OK;
}

int solar_compute_rhs(solar_ctx* c, const double* q, double* q_rhs, int flags) {
  if (!c || !q || !q_rhs) return fail(SOLAR_ERR_ARG, "solar_compute_rhs: null argument");
  const size_t nq = c->nf * 30;
  // q[1] and qw are free between steps; the state in q[0] is not touched.
  int rc = upload_state(c, c->q[1], q, nq);
  if (rc) return rc;
  HIPCHK(solar::launch_gather(c->dt, c->q[1], c->f32, true, 4, 1, c->gxp, c->gyp, c->gxv, c->gyv, c->stream));
  const int mode = 4 | (flags & (SOLAR_RHS_BC | SOLAR_RHS_SPONGE));
  HIPCHK(solar::launch_stage(c->dt, c->q[1], c->qw, c->qw, c->f32, c->gxp, c->gyp, c->gxv, c->gyv, -1, 0.0, mode,
                             c->stream, c->shape, c->c2f));
  return download_state(c, q_rhs, c->qw, nq);
}

int solar_filter_exp(solar_ctx* c, const double* q, double* qf) {
  if (!c || !q || !qf) return fail(SOLAR_ERR_ARG, "solar_filter_exp: null argument");
  const size_t nq = c->nf * 30;
  int rc = upload_state(c, c->q[1], q, nq);
  if (rc) return rc;
  HIPCHK(solar::launch_gather(c->dt, c->q[1], c->f32, false, 0, 5, c->fgx, c->fgy, nullptr, nullptr, c->stream));
  HIPCHK(solar::launch_filter(c->dt, c->q[1], c->qw, c->f32, c->fgx, c->fgy, c->stream, c->shape));
  return download_state(c, qf, c->qw, nq);
}

int solar_compute_rhs_noalloc(solar_ctx* c, const void* q, void* q_rhs, int flags, void* stream) {
  if (!c || !q || !q_rhs) return fail(SOLAR_ERR_ARG, "solar_compute_rhs_noalloc: null argument");
  if (c->scratch_status != hipSuccess || !c->scratch)
    return fail(SOLAR_ERR_HIP, std::string("solar_compute_rhs_noalloc: device scratch unavailable: ") +
                                   hipGetErrorString(c->scratch_status));
  // four ghost sets (p' and velocity, DDX and DDY), each 6 panels x 2 sides x 3 layers x n x nr
  const size_t ng = static_cast<size_t>(12) * c->nr * c->n * 3;
  const size_t set = (ng * sizeof(double) + 255) / 256 * 256;
  if (4 * set > solar::scratch_bytes())
    return fail(SOLAR_ERR_ARG, "solar_compute_rhs_noalloc: n*nr = " + std::to_string(c->n * c->nr) +
                                   " needs " + std::to_string(4 * set) + " bytes of scratch, above the " +
                                   std::to_string(solar::scratch_bytes()) + " available");
  char* base = static_cast<char*>(c->scratch);
  double* gxp = reinterpret_cast<double*>(base);
  double* gyp = reinterpret_cast<double*>(base + set);
  double* gxv = reinterpret_cast<double*>(base + 2 * set);
  double* gyv = reinterpret_cast<double*>(base + 3 * set);
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  HIPCHK(solar::launch_gather(c->dt, q, c->f32, true, 4, 1, gxp, gyp, gxv, gyv, st));
  const int mode = 4 | (flags & (SOLAR_RHS_BC | SOLAR_RHS_SPONGE));
  HIPCHK(solar::launch_stage(c->dt, q, q_rhs, q_rhs, c->f32, gxp, gyp, gxv, gyv, -1, 0.0, mode, st, c->shape,
                             c->c2f));
  return SOLAR_OK;
}

static int ghosts_common(solar_ctx* c, const double* f0, const double* f1, double* gx, double* gy, bool vec) {
  const size_t nf6 = c->nf * 6;
  const size_t ng = static_cast<size_t>(12) * c->nr * c->n * 3;
  // Stage the panel fields into q[1] at the field slots the gather reads,
  // using qw (at least 6*nf doubles in either precision) as the upload buffer.
  double* up = static_cast<double*>(c->qw);
  HIPCHK(hipMemcpyAsync(up, f0, nf6 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(solar::launch_place(up, c->q[1], c->f32, c->nf, vec ? 1 : 4, c->stream));
  if (vec) {
    HIPCHK(hipMemcpyAsync(up, f1, nf6 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(solar::launch_place(up, c->q[1], c->f32, c->nf, 2, c->stream));
    // the scalar slot (field 4) feeds the unused scalar output
    HIPCHK(solar::launch_place(up, c->q[1], c->f32, c->nf, 4, c->stream));
  }
  HIPCHK(solar::launch_gather(c->dt, c->q[1], c->f32, true, 4, 1, c->gxp, c->gyp, c->gxv, c->gyv, c->stream));
  HIPCHK(hipMemcpyAsync(gx, vec ? c->gxv : c->gxp, ng * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(gy, vec ? c->gyv : c->gyp, ng * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SOLAR_OK;
}

int solar_scalar_ghosts(solar_ctx* c, const double* f, double* gx, double* gy) {
  if (!c || !f || !gx || !gy) return fail(SOLAR_ERR_ARG, "solar_scalar_ghosts: null argument");
  // vector slots must hold finite values: fill them with f as well
  double* up = static_cast<double*>(c->qw);
  HIPCHK(hipMemcpyAsync(up, f, c->nf * 6 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(solar::launch_place(up, c->q[1], c->f32, c->nf, 1, c->stream));
  HIPCHK(solar::launch_place(up, c->q[1], c->f32, c->nf, 2, c->stream));
  return ghosts_common(c, f, nullptr, gx, gy, false);
}

int solar_vector_ghosts(solar_ctx* c, const double* vxi, const double* veta, double* gx, double* gy) {
  if (!c || !vxi || !veta || !gx || !gy) return fail(SOLAR_ERR_ARG, "solar_vector_ghosts: null argument");
  return ghosts_common(c, vxi, veta, gx, gy, true);
}

int solar_check_finite(solar_ctx* c) {
  if (!c) return fail(SOLAR_ERR_ARG, "solar_check_finite: null handle");
  HIPCHK(solar::launch_scan(c->q[0], c->f32, c->nf, c->acc, 0, c->stream));
  double h[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(h, c->acc, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (h[0] != 0.0)
    return fail(SOLAR_ERR_NONFINITE, "solar_check_finite: " + std::to_string(static_cast<long long>(h[0])) +
                                         " non-finite values in the state");
  return SOLAR_OK;
}

int solar_norm_vr(solar_ctx* c, double* rms) {
  if (!c || !rms) return fail(SOLAR_ERR_ARG, "solar_norm_vr: null argument");
  HIPCHK(solar::launch_scan(c->q[0], c->f32, c->nf, c->acc, 1, c->stream));
  double h[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(h, c->acc, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *rms = std::sqrt(h[1] / static_cast<double>(c->nf * 6));
  return SOLAR_OK;
}

int solar_write_restart(solar_ctx* c, const char* dir, int cue, double time, int step) {
  if (!c) return fail(SOLAR_ERR_ARG, "solar_write_restart: null handle");
  std::vector<double> h(c->nf * 30);
  int rc = download_state(c, h.data(), c->q[0], h.size());
  return rc ? rc : solar_restart_write(dir, c->n, c->nr, h.data(), cue, time, step);
}

int solar_read_restart(solar_ctx* c, const char* dir, int cue, double* time, int* step) {
  if (!c) return fail(SOLAR_ERR_ARG, "solar_read_restart: null handle");
  std::vector<double> h(c->nf * 30);
  int rc = solar_restart_read(dir, c->n, c->nr, h.data(), cue, time, step);
  if (!rc) rc = upload_state(c, c->q[0], h.data(), h.size());
  if (!rc) HIPCHK(hipStreamSynchronize(c->stream));  // h is released on return
  return rc;
}

int solar_write_visual(solar_ctx* c, const char* dir, int cue, int o_rad) {
  if (!c) return fail(SOLAR_ERR_ARG, "solar_write_visual: null handle");
  std::vector<double> h(c->nf * 30);
  int rc = download_state(c, h.data(), c->q[0], h.size());
  return rc ? rc : solar_visual_write(dir, c->n, c->nr, h.data(), cue, o_rad);
}

int solar_write_radial_trace(solar_ctx* c, const char* dir) {
  if (!c) return fail(SOLAR_ERR_ARG, "solar_write_radial_trace: null handle");
  std::vector<double> h(c->nf * 30);
  int rc = download_state(c, h.data(), c->q[0], h.size());
  return rc ? rc : solar_radial_trace_write(dir, c->n, c->nr, h.data());
}

int solar_norm(solar_ctx* c, int field, int reference_formula, double* out) {
  if (!c || !out || field < 0 || field > 4) return fail(SOLAR_ERR_ARG, "solar_norm: invalid argument");
  HIPCHK(solar::launch_scan(c->q[0], c->f32, c->nf, c->acc, 1, c->stream, field));
  double h[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(h, c->acc, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const double per_panel = static_cast<double>(c->n) * c->n * c->nr;
  // auxiliary.f90:122: (sum / (nx*ny*nr) * 6)**0.5 ; the RMS divides by the 6 panels instead
  *out = reference_formula ? std::pow((h[1] / per_panel) * 6.0, 0.5) : std::sqrt(h[1] / (per_panel * 6.0));
  return SOLAR_OK;
}

int solar_kernel_times(solar_ctx* c, double* out) {
  if (!c || !out) return fail(SOLAR_ERR_ARG, "solar_kernel_times: null argument");
  if (!c->profile) return fail(SOLAR_ERR_ARG, "solar_kernel_times: set SOLAR_PROFILE=1 before solar_init");
  for (int q = 0; q < 5; ++q) out[q] = c->ktime[q];
  return SOLAR_OK;
}

int solar_algorithmic_bytes(solar_ctx* c, double* out) {
  if (!c || !out) return fail(SOLAR_ERR_ARG, "solar_algorithmic_bytes: null argument");
  // in states of 5 fields: reads + writes of q and qw per stage, then the filter
  static const int kStoreAll[6] = {3, 4, 4, 4, 3, 2}, kStoredQ[6] = {2, 4, 3, 4, 3, 2};
  const int* u = c->qw2 ? kStoredQ : kStoreAll;
  for (int x = 0; x < 6; ++x) out[x] = static_cast<double>(u[x]) * 5.0 * static_cast<double>(c->esz);
  return SOLAR_OK;
}

int solar_device_bytes(solar_ctx* c, size_t* bytes) {
  if (!c || !bytes) return fail(SOLAR_ERR_ARG, "solar_device_bytes: null argument");
  *bytes = c->bytes;
  return SOLAR_OK;
}

int solar_host_tables(int n, int nr, const double* gr, const double* c2, const double* rho0, const double* g0,
                      int sponge_len, double a_sponge, double* scal, double* n1, double* r1, double* n2,
                      double* interp, double* ghost, double* sigma, double* drho0dr) {
  Tables t;
  std::string err;
  if (!solar::build_tables(n, nr, gr, c2, rho0, g0, sponge_len, a_sponge, &t, &err))
    return fail(SOLAR_ERR_ARG, "solar_host_tables: " + err);
  scal[0] = t.dxi;
  scal[1] = t.deta;
  scal[2] = t.dchi;
  scal[3] = t.pi;
  const std::vector<double>* a1[8] = {&t.gxi, &t.geta, &t.tanxi, &t.taneta, &t.metric_c, &t.metric_d, &t.global_c,
                                      &t.global_d};
  for (int a = 0; a < 8; ++a) std::memcpy(n1 + static_cast<size_t>(a) * n, a1[a]->data(), sizeof(double) * n);
  std::memcpy(r1, t.gchi.data(), sizeof(double) * nr);
  std::memcpy(r1 + nr, t.drdchi.data(), sizeof(double) * nr);
  std::memcpy(r1 + 2 * nr, t.gr.data(), sizeof(double) * nr);
  const std::vector<double>* a2[4] = {&t.mdelta, &t.sqrtd, &t.xy, &t.cd};
  const size_t n2s = static_cast<size_t>(n) * n;
  for (int a = 0; a < 4; ++a) std::memcpy(n2 + a * n2s, a2[a]->data(), sizeof(double) * n2s);
  std::memcpy(interp, t.interpolant.data(), sizeof(double) * 27 * n);
  const double* g[8] = {t.xm, t.xp, t.ym, t.yp, t.cm, t.cp, t.dm, t.dp};
  for (int a = 0; a < 8; ++a)
    for (int l = 1; l <= 3; ++l) ghost[a * 3 + l - 1] = g[a][l];
  for (int k = 0; k < nr; ++k) {
    sigma[k] = t.radial[k].sigma;
    drho0dr[k] = t.radial[k].drho0dr;
  }
  return SOLAR_OK;
}


// ------------------------------------------------------------ sensitivities

static int adj_alloc(solar_ctx* c) {
  // the adjoint kernels tile the panel exactly (16x16 or 32x8 tiles)
  if (c->n % 16 != 0)
    return fail(SOLAR_ERR_ARG, "adjoint: needs a panel size n that is a multiple of 16 (n = " + std::to_string(c->n) +
                                   "); the forward path takes any n");
  if (c->nr < 12) return fail(SOLAR_ERR_ARG, "adjoint: needs nr >= 12 (the 2-D case nr = 1 runs forward only)");
  if (c->lam) return SOLAR_OK;
  if (c->f32) return fail(SOLAR_ERR_ARG, "adjoint: needs a 64-bit state handle");
  const size_t nq = c->nf * 30;
  int rc;
  if ((rc = dev_alloc(c, &c->wb[0], nq)) || (rc = dev_alloc(c, &c->wb[1], nq)) || (rc = dev_alloc(c, &c->tmp, nq)) ||
      (rc = dev_alloc(c, &c->lam, nq)) || (rc = dev_alloc(c, &c->gc2, c->nf * 6)))
    return rc;
  return SOLAR_OK;
}

static int upload_f64(solar_ctx* c, double* dst, const double* src, size_t count) {
  HIPCHK(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return SOLAR_OK;
}
static int download_f64(solar_ctx* c, double* dst, const double* src, size_t count) {
  HIPCHK(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SOLAR_OK;
}

// lam <- S^T lam for one step: FILTER_EXP^T into tmp, then the five reverse
// stages in place on lam.  With divp, the c^2 gradient of this step is added
// to c->gc2 (divp[s] = div(v) of the input of forward stage s, 0-based).
static int enqueue_adj_step(solar_ctx* c, double dt, double* lam, double* const* divp) {
  const solar::DevTables& T = c->dt;
  HIPCHK(solar::launch_adj_filter(T, lam, c->tmp, c->fgx, c->fgy, c->stream, c->shape.adj_filter_chunks));
  HIPCHK(solar::launch_adj_scatter(T, false, c->fgx, c->fgy, nullptr, nullptr, c->tmp, nullptr, 0.0, 0, 5, c->stream,
                                   true));
  for (int s = 5; s >= 1; --s) {
    solar::AdjStageArgs a;
    // stage 5 starts from Ybar^(5) = F^T lam alone: W^(5) = beta_5 Ybar^(5)
    a.W = s == 5 ? nullptr : c->wb[(4 - s) % 2];
    a.wscale = solar::kBetaRk4[4];
    a.Yin = s == 5 ? c->tmp : lam;
    a.Yout = lam;
    a.Wout = s > 1 ? c->wb[(5 - s) % 2] : nullptr;
    a.dt = dt;
    a.alpha = solar::kAlphaRk4[s - 1];
    a.beta = s > 1 ? solar::kBetaRk4[s - 2] : 0.0;
    a.mode = (s > 1 ? solar::kAdjWriteW : 0) | (divp ? solar::kAdjGrad : 0);
    a.divs = divp ? divp[s - 1] : nullptr;
    a.div32 = c->div_bits == 32;
    a.gc2 = c->gc2;
    a.c2f = c->c2f;
    a.gxa_p = c->gxp;
    a.gya_p = c->gyp;
    a.gxa_v = c->gxv;
    a.gya_v = c->gyv;
    a.chunks = c->shape.adj_stage_chunks;
    HIPCHK(solar::launch_adj_stage(T, a, c->stream));
    HIPCHK(solar::launch_adj_scatter(T, true, c->gxp, c->gyp, c->gxv, c->gyv, lam, a.Wout, a.beta, 4, 1, c->stream));
  }
  return SOLAR_OK;
}

int solar_set_adjoint_options(solar_ctx* c, int div_bits) {
  if (!c || (div_bits != 64 && div_bits != 32)) return fail(SOLAR_ERR_ARG, "solar_set_adjoint_options: div_bits 64 or 32");
  c->div_bits = div_bits;
  return SOLAR_OK;
}

int solar_set_adjoint_cache(solar_ctx* c, int enable) {
  if (!c) return fail(SOLAR_ERR_ARG, "solar_set_adjoint_cache: null handle");
  c->div_cache = enable != 0;
  return SOLAR_OK;
}

int solar_adjoint_recomputed_steps(solar_ctx* c, int* steps) {
  if (!c || !steps) return fail(SOLAR_ERR_ARG, "solar_adjoint_recomputed_steps: null argument");
  *steps = c->adj_recomputed;
  return SOLAR_OK;
}

int solar_set_c2_field(solar_ctx* c, const double* c2) {
  if (!c) return fail(SOLAR_ERR_ARG, "solar_set_c2_field: null handle");
  if (!c2) {
    c->c2f = nullptr;
    return SOLAR_OK;
  }
  double* buf = c->c2f_buf;
  if (!buf) {
    int rc = dev_alloc(c, &buf, c->nf * 6);
    if (rc) return rc;
    c->c2f_buf = buf;
  }
  HIPCHK(hipMemcpyAsync(buf, c2, c->nf * 6 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->c2f = buf;
  return SOLAR_OK;
}

int solar_set_receivers(solar_ctx* c, int nrec, const int* pkji) {
  if (!c || nrec < 0 || (nrec > 0 && !pkji)) return fail(SOLAR_ERR_ARG, "solar_set_receivers: invalid argument");
  std::vector<long long> idx(nrec);
  for (int r = 0; r < nrec; ++r) {
    const int p = pkji[4 * r], k = pkji[4 * r + 1], j = pkji[4 * r + 2], i = pkji[4 * r + 3];
    if (p < 0 || p >= 6 || k < 0 || k >= c->nr || j < 0 || j >= c->n || i < 0 || i >= c->n)
      return fail(SOLAR_ERR_ARG, "solar_set_receivers: receiver " + std::to_string(r) + " outside the grid");
    idx[r] = ((static_cast<long long>(p) * 5 + 4) * c->nr + k) * c->n * c->n + static_cast<long long>(j) * c->n + i;
  }
  if (nrec > c->rec_cap) {
    // the old, smaller buffer goes back before the larger one is taken
    HIPCHK(hipStreamSynchronize(c->stream));
    dev_free(c, c->rec, static_cast<size_t>(c->rec_cap) * sizeof(long long));
    c->rec = nullptr;
    c->rec_cap = 0;
    long long* d = nullptr;
    int rc = dev_alloc(c, &d, static_cast<size_t>(nrec));
    if (rc) return rc;
    c->rec = d;
    c->rec_cap = nrec;
  }
  c->nrec = nrec;
  if (nrec) HIPCHK(hipMemcpyAsync(c->rec, idx.data(), nrec * sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SOLAR_OK;
}

int solar_record(solar_ctx* c, int nsteps, double dt, double* seis) {
  if (!c || nsteps < 0 || (nsteps && !seis)) return fail(SOLAR_ERR_ARG, "solar_record: invalid argument");
  if (c->f32) return fail(SOLAR_ERR_ARG, "solar_record: needs a 64-bit state handle");
  const size_t count = static_cast<size_t>(nsteps) * c->nrec;
  if (count == 0) return solar_rk4(c, nsteps, dt, nullptr);
  void* v = nullptr;
  hipError_t e = hipMalloc(&v, count * sizeof(double));
  if (e != hipSuccess) return fail(SOLAR_ERR_HIP, std::string("solar_record: hipMalloc: ") + hipGetErrorString(e));
  double* d = static_cast<double*>(v);
  int rc = SOLAR_OK;
  for (int s = 0; s < nsteps && rc == SOLAR_OK; ++s) {
    rc = solar_rk4_async(c, 1, dt);
    if (rc == SOLAR_OK) {
      e = solar::launch_sample(static_cast<const double*>(c->q[0]), c->rec, c->nrec, d + (size_t)s * c->nrec,
                               c->stream);
      if (e != hipSuccess) rc = fail(SOLAR_ERR_HIP, std::string("solar_record: ") + hipGetErrorString(e));
    }
  }
  if (rc == SOLAR_OK) rc = download_f64(c, seis, d, count);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(v);
  return rc;
}

int solar_tangent_linear(solar_ctx* c, const double* q0, int nsteps, double dt, const double* dc2, double* dseis) {
  if (!c || !q0 || !dc2 || !dseis || nsteps < 1) return fail(SOLAR_ERR_ARG, "solar_tangent_linear: invalid argument");
  if (c->nrec < 1) return fail(SOLAR_ERR_ARG, "solar_tangent_linear: no receivers set");
  int rc = adj_alloc(c);
  if (rc) return rc;
  const size_t nq = c->nf * 30, nf6 = c->nf * 6, nrec = static_cast<size_t>(c->nrec);
  const size_t nseis = static_cast<size_t>(nsteps) * nrec;
  void* v = nullptr;
  hipError_t e = hipMalloc(&v, (5 * nf6 + nf6 + nseis) * sizeof(double));
  if (e != hipSuccess) return fail(SOLAR_ERR_HIP, std::string("solar_tangent_linear: hipMalloc: ") + hipGetErrorString(e));
  double* divb = static_cast<double*>(v);
  double* dc2d = divb + 5 * nf6;
  double* ds = dc2d + nf6;
  double* dp[5];
  for (int s = 0; s < 5; ++s) dp[s] = divb + s * nf6;
  const int bits = c->div_bits;
  c->div_bits = 64;  // the tangent-linear source takes FP64 primal divergences
  auto run = [&]() -> int {
    int r;
    if ((r = upload_state(c, c->q[0], q0, nq))) return r;
    if ((r = upload_f64(c, dc2d, dc2, nf6))) return r;
    HIPCHK(hipMemsetAsync(c->wb[0], 0, nq * sizeof(double), c->stream));
    for (int n = 0; n < nsteps; ++n) {
      if ((r = enqueue_step(c, dt, false, dp))) return r;  // primal, keeping its stage divergences
      if ((r = enqueue_step(c, dt, false, nullptr, c->wb[0], c->wb[1], c->tmp, dc2d, dp))) return r;
      HIPCHK(solar::launch_sample(c->wb[0], c->rec, c->nrec, ds + n * nrec, c->stream));
    }
    return download_f64(c, dseis, ds, nseis);
  };
  rc = run();
  (void)hipStreamSynchronize(c->stream);
  c->div_bits = bits;
  (void)hipFree(v);
  return rc;
}

int solar_adj_rhs(solar_ctx* c, const double* w, double* out) {
  if (!c || !w || !out) return fail(SOLAR_ERR_ARG, "solar_adj_rhs: null argument");
  int rc = adj_alloc(c);
  if (rc) return rc;
  const size_t nq = c->nf * 30;
  if ((rc = upload_f64(c, c->wb[0], w, nq))) return rc;
  solar::AdjStageArgs a;
  a.W = c->wb[0];
  a.Yout = c->tmp;
  a.c2f = c->c2f;
  a.mode = solar::kAdjRhsOnly;
  a.gxa_p = c->gxp;
  a.gya_p = c->gyp;
  a.gxa_v = c->gxv;
  a.gya_v = c->gyv;
  a.chunks = c->shape.adj_stage_chunks;
  HIPCHK(solar::launch_adj_stage(c->dt, a, c->stream));
  HIPCHK(solar::launch_adj_scatter(c->dt, true, c->gxp, c->gyp, c->gxv, c->gyv, c->tmp, nullptr, 0.0, 4, 1, c->stream));
  return download_f64(c, out, c->tmp, nq);
}

int solar_adj_filter(solar_ctx* c, const double* o, double* out) {
  if (!c || !o || !out) return fail(SOLAR_ERR_ARG, "solar_adj_filter: null argument");
  int rc = adj_alloc(c);
  if (rc) return rc;
  const size_t nq = c->nf * 30;
  if ((rc = upload_f64(c, c->lam, o, nq))) return rc;
  HIPCHK(solar::launch_adj_filter(c->dt, c->lam, c->tmp, c->fgx, c->fgy, c->stream, c->shape.adj_filter_chunks));
  HIPCHK(solar::launch_adj_scatter(c->dt, false, c->fgx, c->fgy, nullptr, nullptr, c->tmp, nullptr, 0.0, 0, 5,
                                   c->stream, true));
  return download_f64(c, out, c->tmp, nq);
}

int solar_adj_step(solar_ctx* c, double dt, const double* lam, double* out) {
  if (!c || !lam || !out) return fail(SOLAR_ERR_ARG, "solar_adj_step: null argument");
  int rc = adj_alloc(c);
  if (rc) return rc;
  const size_t nq = c->nf * 30;
  if ((rc = upload_f64(c, c->lam, lam, nq))) return rc;
  if ((rc = enqueue_adj_step(c, dt, c->lam, nullptr))) return rc;
  return download_f64(c, out, c->lam, nq);
}

int solar_misfit_gradient(solar_ctx* c, const double* q0, int nsteps, double dt, const double* data,
                          int checkpoint_every, double* J, double* grad, double* seis_out, int* segment_used) {
  if (!c || !q0 || !data || !J || !grad || nsteps < 1)
    return fail(SOLAR_ERR_ARG, "solar_misfit_gradient: invalid argument");
  if (c->nrec < 1) return fail(SOLAR_ERR_ARG, "solar_misfit_gradient: no receivers set");
  int rc = adj_alloc(c);
  if (rc) return rc;
  const size_t nq = c->nf * 30, nf6 = c->nf * 6;
  const size_t esz_div = c->div_bits == 32 ? sizeof(float) : sizeof(double);
  const size_t step_div = 5 * nf6 * esz_div, ck_bytes = nq * sizeof(double);
  const size_t trace_bytes = 2 * static_cast<size_t>(nsteps) * c->nrec * sizeof(double);
  // Segment length L: checkpoint_every, or (<= 0) the longest that fits.  The
  // last L steps keep their stage divergences from the forward pass; every
  // earlier segment is recomputed once from its checkpoint.
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };  // 256-byte aligned parts of one workspace
  auto need = [&](int L) {
    const size_t nseg = (static_cast<size_t>(nsteps) + L - 1) / L;
    return up(static_cast<size_t>(L) * step_div) + up((nseg - 1) * ck_bytes) + 2 * up(trace_bytes / 2);
  };
  size_t freeb = 0, totalb = 0;
  HIPCHK(hipMemGetInfo(&freeb, &totalb));
  freeb += c->adj_ws_bytes;  // the handle's workspace is reused or replaced
  const size_t margin = size_t(2) << 30;
  const size_t avail = freeb > margin ? freeb - margin : 0;
  int L = checkpoint_every > 0 ? std::min(checkpoint_every, nsteps) : nsteps;
  if (checkpoint_every <= 0) {  // the longest segment that fits (no cache needed beyond it)
    while (L > 1 && need(L) > avail) --L;
    // keep the handle's workspace when it holds nearly as long a segment (a
    // new allocation of hundreds of GB costs seconds)
    if (c->adj_ws_bytes && need(L) > c->adj_ws_bytes) {
      int Lw = L;
      while (Lw > 1 && need(Lw) > c->adj_ws_bytes) --Lw;
      if (need(Lw) <= c->adj_ws_bytes && Lw * 10 >= L * 9) L = Lw;
    }
  }
  if (need(L) > avail)
    return fail(SOLAR_ERR_ARG, "solar_misfit_gradient: needs " + std::to_string(need(L) >> 20) + " MiB, " +
                                   std::to_string(avail >> 20) + " MiB free" +
                                   (checkpoint_every > 0 ? "; lower checkpoint_every" : ""));
  const int nseg = (nsteps + L - 1) / L;
  std::vector<int> n0(nseg), n1(nseg);  // segment g covers steps [n0, n1); the last is full
  for (int g = 0; g < nseg; ++g) {
    n1[g] = nsteps - (nseg - 1 - g) * L;
    n0[g] = std::max(0, n1[g] - L);
  }
  // Divergence cache: with a given checkpoint spacing, free memory beyond
  // the one-segment buffer keeps the divergences of the last C + 1 segments
  // from the forward pass, so that only the earlier segments are recomputed.
  int C = 0;
  if (checkpoint_every > 0 && c->div_cache) {
    auto need_c = [&](int cc) {
      return need(L) - up(static_cast<size_t>(L) * step_div) + up(static_cast<size_t>(L) * (cc + 1) * step_div);
    };
    while (C + 1 < nseg && need_c(C + 1) <= avail) ++C;
  }
  const int gkeep = nseg - 1 - C;      // segments g >= gkeep keep their divergences from the forward pass
  const int nkeep = n0[gkeep];         // first step whose divergences the forward pass stores
  const int D = nsteps - nkeep;        // steps of divergences held (at least L)
  // one workspace, kept by the handle for the next call: checkpoints | divergences | traces | data
  const size_t b_ck = up((nseg - 1) * ck_bytes), b_div = up(static_cast<size_t>(std::max(D, L)) * step_div);
  const size_t b_tr = up(trace_bytes / 2), ws_need = b_ck + b_div + 2 * b_tr;
  const auto t_start = std::chrono::steady_clock::now();
  if (ws_need > c->adj_ws_bytes) {
    if (c->adj_ws) (void)hipFree(c->adj_ws);
    c->adj_ws = nullptr;
    c->adj_ws_bytes = 0;
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, ws_need);
    if (e != hipSuccess) return fail(SOLAR_ERR_HIP, std::string("hipMalloc (gradient workspace): ") + hipGetErrorString(e));
    c->adj_ws = v;
    c->adj_ws_bytes = ws_need;
  }
  char* wsb = static_cast<char*>(c->adj_ws);
  double* ck = reinterpret_cast<double*>(wsb);
  double* divs = reinterpret_cast<double*>(wsb + b_ck);
  double* seis = reinterpret_cast<double*>(wsb + b_ck + b_div);
  double* dat = reinterpret_cast<double*>(wsb + b_ck + b_div + b_tr);
  const auto t_ws = std::chrono::steady_clock::now();
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  for (auto& x : ev) HIPCHK(hipEventCreate(&x));
  float ms_rec = 0.f, ms_adj = 0.f, ms_fwd = 0.f;
  std::vector<double*> dp(static_cast<size_t>(std::max(D, L)) * 5);
  for (size_t x = 0; x < dp.size(); ++x)
    dp[x] = reinterpret_cast<double*>(reinterpret_cast<char*>(divs) + x * nf6 * esz_div);
  const size_t nrec = static_cast<size_t>(c->nrec);
  auto run = [&]() -> int {
    int r;
    if ((r = upload_f64(c, dat, data, static_cast<size_t>(nsteps) * nrec))) return r;
    if ((r = upload_state(c, c->q[0], q0, nq))) return r;
    // forward pass: checkpoints, traces, and the stage divergences of the
    // last segment (of the last C + 1 with the cache)
    const int nlast = nkeep;
    for (int n = 0, g = 0; n < nsteps; ++n) {
      if (g < nseg - 1 && n == n0[g]) {
        HIPCHK(hipMemcpyAsync(ck + static_cast<size_t>(g) * nq, c->q[0], ck_bytes, hipMemcpyDeviceToDevice,
                              c->stream));
        ++g;
      }
      if (n >= nlast) {
        if ((r = enqueue_step(c, dt, false, &dp[static_cast<size_t>(n - nlast) * 5]))) return r;
      } else if ((r = solar_rk4_async(c, 1, dt))) {
        return r;
      }
      HIPCHK(solar::launch_sample(static_cast<const double*>(c->q[0]), c->rec, c->nrec, seis + n * nrec, c->stream));
    }
    std::vector<double> sh(static_cast<size_t>(nsteps) * nrec);
    HIPCHK(hipEventRecord(ev[0], c->stream));
    if ((r = download_f64(c, sh.data(), seis, sh.size()))) return r;
    {
      const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_ws).count();
      ms_fwd = static_cast<float>(s * 1e3);
    }
    double jj = 0.0;
    for (size_t x = 0; x < sh.size(); ++x) jj += (sh[x] - data[x]) * (sh[x] - data[x]);
    *J = jj;
    if (seis_out) std::memcpy(seis_out, sh.data(), sh.size() * sizeof(double));
    // reverse pass
    HIPCHK(hipMemsetAsync(c->lam, 0, nq * sizeof(double), c->stream));
    HIPCHK(hipMemsetAsync(c->gc2, 0, nf6 * sizeof(double), c->stream));
    for (int g = nseg - 1; g >= 0; --g) {
      HIPCHK(hipEventRecord(ev[0], c->stream));
      // divergences of step n at dp[(n - base) * 5]: cached from the forward
      // pass, or recomputed from the segment's checkpoint into the buffer's start
      const int base = g >= gkeep ? nkeep : n0[g];
      if (g < gkeep) {  // recompute the segment from its checkpoint, keeping the divergences
        HIPCHK(hipMemcpyAsync(c->q[0], ck + static_cast<size_t>(g) * nq, ck_bytes, hipMemcpyDeviceToDevice,
                              c->stream));
        for (int m = n0[g]; m < n1[g]; ++m)
          if ((r = enqueue_step(c, dt, false, &dp[static_cast<size_t>(m - base) * 5]))) return r;
      }
      HIPCHK(hipEventRecord(ev[1], c->stream));
      for (int n = n1[g] - 1; n >= n0[g]; --n) {
        HIPCHK(solar::launch_inject(c->lam, c->rec, c->nrec, seis + n * nrec, dat + n * nrec, c->stream));
        if ((r = enqueue_adj_step(c, dt, c->lam, &dp[static_cast<size_t>(n - base) * 5]))) return r;
      }
      HIPCHK(hipEventRecord(ev[2], c->stream));
      HIPCHK(hipEventSynchronize(ev[2]));
      float a = 0.f, b = 0.f;
      HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]));
      HIPCHK(hipEventElapsedTime(&b, ev[1], ev[2]));
      ms_rec += a;
      ms_adj += b;
    }
    return download_f64(c, grad, c->gc2, nf6);
  };
  rc = run();
  (void)hipStreamSynchronize(c->stream);
  for (auto& x : ev) (void)hipEventDestroy(x);
  c->adj_times[0] = std::chrono::duration<double>(t_ws - t_start).count();
  c->adj_times[1] = ms_fwd * 1e-3;
  c->adj_times[2] = ms_rec * 1e-3;
  c->adj_times[3] = ms_adj * 1e-3;
  c->adj_recomputed = nkeep;  // steps recomputed from checkpoints
  if (segment_used) *segment_used = L;
  return rc;
}

int solar_adjoint_times(solar_ctx* c, double* out) {
  if (!c || !out) return fail(SOLAR_ERR_ARG, "solar_adjoint_times: null argument");
  for (int x = 0; x < 4; ++x) out[x] = c->adj_times[x];
  return SOLAR_OK;
}
}  // extern "C"
