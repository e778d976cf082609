// Batched ensemble of independent simulations with uneven grids: every
// member is one cube.f90 run (its own n, nr, background, sponge and dt), and
// one RK4 step of the whole ensemble is 5 x (one gather launch + one stage
// launch per tile class) + one gather + one filter launch per tile class,
// captured once in a hipGraph.  Tile shapes and radial chunk lengths per
// member class come from a timing autotuner (solar_ens_autotune).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/solar_amd.h"
#include "solar_internal.hpp"
#include "solar_kernels.hpp"
#include "solar_tables.hpp"

using solar::DevTables;
using solar::EnsMember;
using solar::Tables;

namespace {

#define ENSCHK(expr)                                                                                 \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) return solar::set_error(SOLAR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Launch parameters of one member: stage tile, stage chunk, filter tile, filter chunk.
struct Choice {
  int sti = 0, stj = 0, skc = 0, fti = 0, ftj = 0, fkc = 0;
};

struct Launch {
  int ti = 0, tj = 0, nitems = 0, grid = 0;
  int* items = nullptr;  // device, 4 ints per item
};

int xcd_perm(int b, int nb) { return (nb % 8 != 0) ? b : (b % 8) * (nb / 8) + b / 8; }

}  // namespace

struct solar_ens {
  int nm = 0;
  std::vector<int> n, nr;
  std::vector<EnsMember> mem;  // host copy of the member table
  EnsMember* dmem = nullptr;
  std::vector<Choice> choice;
  std::vector<Launch> stage_l, filter_l;
  // gathers: one launch per panel size n (its block size is 3n strip
  // points); table = first block of each member of the group, then the
  // members' indices
  struct GatherGroup {
    int* table = nullptr;
    int nm = 0, nblocks = 0, n = 0;
  };
  std::vector<GatherGroup> gg;
  int gchunks = 4;  // radial chunks per receiver edge in the gathers
  int persist = 0;   // > 0: grid cap of the work-list launches
  std::vector<void*> allocs;
  std::vector<void*> list_allocs;  // work lists, rebuilt by the autotuner
  size_t bytes = 0;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

int ens_alloc(solar_ens* e, void** p, size_t b, std::vector<void*>* list) {
  hipError_t he = hipMalloc(p, b ? b : 8);
  if (he != hipSuccess)
    return solar::set_error(SOLAR_ERR_HIP, "hipMalloc(" + std::to_string(b) + " bytes): " + hipGetErrorString(he));
  list->push_back(*p);
  e->bytes += b;
  return SOLAR_OK;
}

void drop_graph(solar_ens* e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
}

void release(solar_ens* e) {
  drop_graph(e);
  for (void* p : e->list_allocs) (void)hipFree(p);
  for (void* p : e->allocs) (void)hipFree(p);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  e->list_allocs.clear();
  e->allocs.clear();
}

bool stage_shape_ok(int n, int ti, int tj) {
  const bool known = (ti == 64 && tj == 8) || (ti == 32 && tj == 8) || (ti == 16 && tj == 16);
  return known && n % ti == 0 && n % tj == 0;
}
bool filter_shape_ok(int n, int ti, int tj) {
  const bool known = (ti == 32 && tj == 16) || (ti == 16 && tj == 16) || (ti == 32 && tj == 8);
  return known && n % ti == 0 && n % tj == 0;
}

Choice default_choice(int n, int nr) {
  Choice c;
  c.sti = n % 32 == 0 ? 32 : 16;
  c.stj = c.sti == 16 ? 16 : 8;
  c.skc = std::min(nr, 32);
  c.fti = n % 32 == 0 ? 32 : 16;
  c.ftj = 16;
  c.fkc = std::min(nr, 48);
  return c;
}

// Work items of member m for the stage (panels) or the filter (5P+F).
void member_items(const solar_ens* e, int m, bool filter, int ti, int tj, int kc, std::vector<int>* out) {
  const int n = e->n[m], nr = e->nr[m];
  const int nti = n / ti, ntj = n / tj, npf = filter ? 30 : 6;
  for (int k0 = 0; k0 < nr; k0 += kc)
    for (int p = 0; p < npf; ++p)
      for (int bj = 0; bj < ntj; ++bj)
        for (int bi = 0; bi < nti; ++bi) {
          out->push_back(m);
          out->push_back(bi | (bj << 16));
          out->push_back(p | (k0 << 8));
          out->push_back(std::min(k0 + kc, nr));
        }
}

// Uploads one launch's items in dispatch order (neighbouring logical items on one XCD).
int make_launch(solar_ens* e, int ti, int tj, const std::vector<int>& logical, Launch* L) {
  const int ni = static_cast<int>(logical.size() / 4);
  std::vector<int> disp(logical.size());
  for (int b = 0; b < ni; ++b) std::memcpy(&disp[4 * b], &logical[4 * xcd_perm(b, ni)], 4 * sizeof(int));
  void* d = nullptr;
  int rc = ens_alloc(e, &d, disp.size() * sizeof(int), &e->list_allocs);
  if (rc) return rc;
  ENSCHK(hipMemcpy(d, disp.data(), disp.size() * sizeof(int), hipMemcpyHostToDevice));
  L->ti = ti;
  L->tj = tj;
  L->nitems = ni;
  L->grid = e->persist > 0 ? std::min(ni, e->persist) : ni;
  L->items = static_cast<int*>(d);
  return SOLAR_OK;
}

// Groups members by tile shape: one stage launch and one filter launch per shape.
int build_lists(solar_ens* e) {
  drop_graph(e);
  for (void* p : e->list_allocs) (void)hipFree(p);
  e->list_allocs.clear();
  e->stage_l.clear();
  e->filter_l.clear();
  std::map<std::pair<int, int>, std::vector<int>> sg, fg;
  for (int m = 0; m < e->nm; ++m) {
    const Choice& c = e->choice[m];
    member_items(e, m, false, c.sti, c.stj, c.skc, &sg[{c.sti, c.stj}]);
    member_items(e, m, true, c.fti, c.ftj, c.fkc, &fg[{c.fti, c.ftj}]);
  }
  for (auto& kv : sg) {
    Launch L;
    int rc = make_launch(e, kv.first.first, kv.first.second, kv.second, &L);
    if (rc) return rc;
    e->stage_l.push_back(L);
  }
  for (auto& kv : fg) {
    Launch L;
    int rc = make_launch(e, kv.first.first, kv.first.second, kv.second, &L);
    if (rc) return rc;
    e->filter_l.push_back(L);
  }
  return SOLAR_OK;
}

// Every stage also writes its output's W/E edge columns to the member's
// ecols, from which the next gather reads its column strips (compact lines
// instead of three values from every row of the state); the step's first
// gather reads the filtered state.
int enqueue_step(solar_ens* e) {
  int cur = 0;
  for (int s = 0; s < 5; ++s) {
    for (const auto& G : e->gg)
      ENSCHK(solar::launch_gather_ens(e->dmem, G.table, G.nm, G.nblocks, G.n, true, cur, s > 0, e->gchunks,
                                      e->stream));
    for (const Launch& L : e->stage_l)
      ENSCHK(solar::launch_stage_ens(e->dmem, L.items, L.nitems, L.grid, L.ti, L.tj, s, cur, e->stream));
    cur ^= 1;
  }
  for (const auto& G : e->gg)
    ENSCHK(solar::launch_gather_ens(e->dmem, G.table, G.nm, G.nblocks, G.n, false, cur, true, e->gchunks,
                                    e->stream));
  for (const Launch& L : e->filter_l)
    ENSCHK(solar::launch_filter_ens(e->dmem, L.items, L.nitems, L.grid, L.ti, L.tj, cur, e->stream));
  return SOLAR_OK;
}

int clear_state(solar_ens* e) {
  for (int m = 0; m < e->nm; ++m) {
    const size_t b = static_cast<size_t>(e->n[m]) * e->n[m] * e->nr[m] * 30 * sizeof(double);
    ENSCHK(hipMemsetAsync(e->mem[m].q[0], 0, b, e->stream));
    ENSCHK(hipMemsetAsync(e->mem[m].q[1], 0, b, e->stream));
    ENSCHK(hipMemsetAsync(e->mem[m].qw, 0, b, e->stream));
  }
  ENSCHK(hipStreamSynchronize(e->stream));
  return SOLAR_OK;
}

// Mean time (ms) of reps launches of `launch` on the ensemble's stream.
template <typename F>
int time_launch(solar_ens* e, int reps, F&& launch, float* ms) {
  hipEvent_t a, b;
  ENSCHK(hipEventCreate(&a));
  ENSCHK(hipEventCreate(&b));
  int rc = SOLAR_OK;
  hipError_t he = launch();  // warm-up
  if (he == hipSuccess) he = hipEventRecord(a, e->stream);
  for (int r = 0; r < reps && he == hipSuccess; ++r) he = launch();
  if (he == hipSuccess) he = hipEventRecord(b, e->stream);
  if (he == hipSuccess) he = hipEventSynchronize(b);
  if (he == hipSuccess) he = hipEventElapsedTime(ms, a, b);
  if (he != hipSuccess) rc = solar::set_error(SOLAR_ERR_HIP, std::string("autotune: ") + hipGetErrorString(he));
  *ms /= static_cast<float>(reps > 0 ? reps : 1);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

}  // namespace

extern "C" {

int solar_ens_create(solar_ens** out, int nmembers, const int* n, const int* nr, const double* gr, const double* c2,
                     const double* rho0, const double* g0, const int* sponge_len, const double* a_sponge,
                     const double* dt) {
  if (!out || nmembers < 1 || !n || !nr || !gr || !c2 || !rho0 || !g0 || !sponge_len || !a_sponge || !dt)
    return solar::set_error(SOLAR_ERR_ARG, "solar_ens_create: invalid argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return solar::set_error(SOLAR_ERR_HIP, "solar_ens_create: no HIP device available");
  solar_ens* e = new solar_ens();
  auto bail = [&](int code) {
    release(e);
    delete e;
    return code;
  };
  e->nm = nmembers;
  e->n.assign(n, n + nmembers);
  e->nr.assign(nr, nr + nmembers);
  e->mem.resize(nmembers);
  e->choice.resize(nmembers);
  if (const char* v = std::getenv("SOLAR_ENS_GATHER_CHUNKS")) {  // launch-shape sweeps
    const int c = std::atoi(v);
    if (c >= 1 && c <= 64) e->gchunks = c;
  }
  if (hipStreamCreate(&e->stream) != hipSuccess) return bail(solar::set_error(SOLAR_ERR_HIP, "hipStreamCreate failed"));
  size_t roff = 0;
  for (int m = 0; m < nmembers; ++m) {
    if (!(dt[m] > 0.0))
      return bail(solar::set_error(SOLAR_ERR_ARG, "solar_ens_create: member " + std::to_string(m) + ": dt <= 0"));
    Tables t;
    std::string err;
    if (n[m] % 16 != 0)  // the ensemble's work lists tile every member exactly
      return bail(solar::set_error(SOLAR_ERR_ARG, "solar_ens_create: member " + std::to_string(m) +
                                                      ": n must be a multiple of 16 in an ensemble"));
    if (nr[m] < 12)
      return bail(solar::set_error(SOLAR_ERR_ARG, "solar_ens_create: member " + std::to_string(m) +
                                                      ": nr must be at least 12 in an ensemble"));
    if (!solar::build_tables(n[m], nr[m], gr + roff, c2 + roff, rho0 + roff, g0 + roff, sponge_len[m], a_sponge[m],
                             &t, &err))
      return bail(solar::set_error(SOLAR_ERR_ARG, "solar_ens_create: member " + std::to_string(m) + ": " + err));
    roff += static_cast<size_t>(nr[m]);
    EnsMember& M = e->mem[m];
    int rc = solar::upload_tables(t, &M.T, &e->allocs, &e->bytes);
    if (rc) return bail(rc);
    const size_t nq = static_cast<size_t>(n[m]) * n[m] * nr[m] * 30;
    const size_t ng = static_cast<size_t>(12) * nr[m] * n[m] * 3;
    void* p[10];
    const size_t sz[10] = {nq, nq, nq, ng, ng, ng, ng, 5 * ng, 5 * ng, solar::edge_cols_elems(n[m], nr[m])};
    for (int a = 0; a < 10; ++a) {
      if ((rc = ens_alloc(e, &p[a], sz[a] * sizeof(double), &e->allocs))) return bail(rc);
      if (hipMemset(p[a], 0, sz[a] * sizeof(double)) != hipSuccess)
        return bail(solar::set_error(SOLAR_ERR_HIP, "solar_ens_create: hipMemset failed"));
    }
    M.q[0] = static_cast<double*>(p[0]);
    M.q[1] = static_cast<double*>(p[1]);
    M.qw = static_cast<double*>(p[2]);
    M.gxp = static_cast<double*>(p[3]);
    M.gyp = static_cast<double*>(p[4]);
    M.gxv = static_cast<double*>(p[5]);
    M.gyv = static_cast<double*>(p[6]);
    M.fgx = static_cast<double*>(p[7]);
    M.fgy = static_cast<double*>(p[8]);
    M.ecols = static_cast<double*>(p[9]);
    M.dt = dt[m];
    e->choice[m] = default_choice(n[m], nr[m]);
  }
  void* d = nullptr;
  int rc;
  if ((rc = ens_alloc(e, &d, nmembers * sizeof(EnsMember), &e->allocs))) return bail(rc);
  e->dmem = static_cast<EnsMember*>(d);
  if (hipMemcpy(e->dmem, e->mem.data(), nmembers * sizeof(EnsMember), hipMemcpyHostToDevice) != hipSuccess)
    return bail(solar::set_error(SOLAR_ERR_HIP, "solar_ens_create: member table upload failed"));
  std::map<int, std::vector<int>> by_n;
  for (int m = 0; m < nmembers; ++m) by_n[n[m]].push_back(m);
  for (auto& kv : by_n) {
    solar_ens::GatherGroup G;
    G.n = kv.first;
    G.nm = static_cast<int>(kv.second.size());
    std::vector<int> table(2 * G.nm);
    for (int x = 0; x < G.nm; ++x) {
      const int m = kv.second[x];
      table[x] = G.nblocks;
      table[G.nm + x] = m;
      G.nblocks += 24 * solar::gather_blocks_per_edge(nr[m], e->gchunks);
    }
    if ((rc = ens_alloc(e, &d, table.size() * sizeof(int), &e->allocs))) return bail(rc);
    G.table = static_cast<int*>(d);
    if (hipMemcpy(G.table, table.data(), table.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      return bail(solar::set_error(SOLAR_ERR_HIP, "solar_ens_create: gather table upload failed"));
    e->gg.push_back(G);
  }
  if ((rc = build_lists(e))) return bail(rc);
  if ((rc = clear_state(e))) return bail(rc);
  *out = e;
  solar::clear_error();
  return SOLAR_OK;
}

int solar_ens_finalize(solar_ens* e) {
  if (!e) return SOLAR_OK;
  (void)hipStreamSynchronize(e->stream);
  release(e);
  delete e;
  return SOLAR_OK;
}

int solar_ens_set_tiles(solar_ens* e, int m, const int* six) {
  if (!e || !six || m < 0 || m >= e->nm) return solar::set_error(SOLAR_ERR_ARG, "solar_ens_set_tiles: invalid argument");
  Choice c{six[0], six[1], six[2], six[3], six[4], six[5]};
  if (!stage_shape_ok(e->n[m], c.sti, c.stj) || !filter_shape_ok(e->n[m], c.fti, c.ftj) || c.skc < 1 || c.fkc < 1)
    return solar::set_error(SOLAR_ERR_ARG, "solar_ens_set_tiles: unsupported tile shape or chunk for member " +
                                               std::to_string(m));
  e->choice[m] = c;
  return build_lists(e);
}

int solar_ens_tiles(solar_ens* e, int m, int* six) {
  if (!e || !six || m < 0 || m >= e->nm) return solar::set_error(SOLAR_ERR_ARG, "solar_ens_tiles: invalid argument");
  const Choice& c = e->choice[m];
  const int v[6] = {c.sti, c.stj, c.skc, c.fti, c.ftj, c.fkc};
  std::memcpy(six, v, sizeof(v));
  return SOLAR_OK;
}

int solar_ens_autotune(solar_ens* e, int reps) {
  if (!e || reps < 1) return solar::set_error(SOLAR_ERR_ARG, "solar_ens_autotune: invalid argument");
  drop_graph(e);
  // classes: members sharing (n, nr) take the same shapes
  std::map<std::pair<int, int>, std::vector<int>> cls;
  for (int m = 0; m < e->nm; ++m) cls[{e->n[m], e->nr[m]}].push_back(m);
  const int sshape[3][2] = {{64, 8}, {32, 8}, {16, 16}};
  const int fshape[3][2] = {{32, 16}, {16, 16}, {32, 8}};
  const int skcs[3] = {16, 32, 64}, fkcs[2] = {48, 96};
  int rc;
  for (auto& kv : cls) {
    const int n = kv.first.first, nr = kv.first.second;
    Choice best = default_choice(n, nr);
    float best_s = 1e30f, best_f = 1e30f;
    for (auto& sh : sshape)
      for (int kc : skcs) {
        if (!stage_shape_ok(n, sh[0], sh[1]) || (kc > nr && kc != skcs[0])) continue;
        std::vector<int> items;
        for (int m : kv.second) member_items(e, m, false, sh[0], sh[1], std::min(kc, nr), &items);
        Launch L;
        if ((rc = make_launch(e, sh[0], sh[1], items, &L))) return rc;
        float ms = 0.f;
        if ((rc = time_launch(e, reps, [&]() {
               return solar::launch_stage_ens(e->dmem, L.items, L.nitems, L.grid, L.ti, L.tj, 2, 0, e->stream);
             }, &ms)))
          return rc;
        if (ms < best_s) {
          best_s = ms;
          best.sti = sh[0];
          best.stj = sh[1];
          best.skc = std::min(kc, nr);
        }
      }
    for (auto& sh : fshape)
      for (int kc : fkcs) {
        if (!filter_shape_ok(n, sh[0], sh[1]) || (kc > nr && kc != fkcs[0])) continue;
        std::vector<int> items;
        for (int m : kv.second) member_items(e, m, true, sh[0], sh[1], std::min(kc, nr), &items);
        Launch L;
        if ((rc = make_launch(e, sh[0], sh[1], items, &L))) return rc;
        float ms = 0.f;
        if ((rc = time_launch(e, reps, [&]() {
               return solar::launch_filter_ens(e->dmem, L.items, L.nitems, L.grid, L.ti, L.tj, 0, e->stream);
             }, &ms)))
          return rc;
        if (ms < best_f) {
          best_f = ms;
          best.fti = sh[0];
          best.ftj = sh[1];
          best.fkc = std::min(kc, nr);
        }
      }
    for (int m : kv.second) e->choice[m] = best;
  }
  // The classes were timed one at a time, but a step launches them merged by
  // shape: the tuned choices are kept only if the whole step beats the
  // default shapes, which merge into the fewest launches.
  auto step_ms = [&](float* ms) {
    return time_launch(e, reps, [&]() { return enqueue_step(e) ? hipErrorUnknown : hipSuccess; }, ms);
  };
  e->persist = 0;
  float t_tuned = 0.f, t_default = 0.f;
  if ((rc = build_lists(e)) || (rc = step_ms(&t_tuned))) return rc;
  std::vector<Choice> tuned(e->choice.begin(), e->choice.end());
  for (int m = 0; m < e->nm; ++m) e->choice[m] = default_choice(e->n[m], e->nr[m]);
  if ((rc = build_lists(e)) || (rc = step_ms(&t_default))) return rc;
  if (t_tuned < t_default) std::copy(tuned.begin(), tuned.end(), e->choice.begin());
  // then the grid policy: one block per item, or a persistent grid
  float t_items = std::min(t_tuned, t_default), t_pers = 0.f;
  if ((rc = build_lists(e))) return rc;
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  e->persist = 4 * (cus > 0 ? cus : 256);
  if ((rc = build_lists(e)) || (rc = step_ms(&t_pers))) return rc;
  if (t_items <= t_pers) {
    e->persist = 0;
    if ((rc = build_lists(e))) return rc;
  }
  return clear_state(e);
}

int solar_ens_set_state(solar_ens* e, int m, const double* q) {
  if (!e || !q || m < 0 || m >= e->nm) return solar::set_error(SOLAR_ERR_ARG, "solar_ens_set_state: invalid argument");
  const size_t b = static_cast<size_t>(e->n[m]) * e->n[m] * e->nr[m] * 30 * sizeof(double);
  ENSCHK(hipMemcpyAsync(e->mem[m].q[0], q, b, hipMemcpyHostToDevice, e->stream));
  ENSCHK(hipStreamSynchronize(e->stream));
  return SOLAR_OK;
}

int solar_ens_get_state(solar_ens* e, int m, double* q) {
  if (!e || !q || m < 0 || m >= e->nm) return solar::set_error(SOLAR_ERR_ARG, "solar_ens_get_state: invalid argument");
  const size_t b = static_cast<size_t>(e->n[m]) * e->n[m] * e->nr[m] * 30 * sizeof(double);
  ENSCHK(hipMemcpyAsync(q, e->mem[m].q[0], b, hipMemcpyDeviceToHost, e->stream));
  ENSCHK(hipStreamSynchronize(e->stream));
  return SOLAR_OK;
}

int solar_ens_rk4_async(solar_ens* e, int nsteps) {
  if (!e || nsteps < 0) return solar::set_error(SOLAR_ERR_ARG, "solar_ens_rk4: invalid argument");
  if (!e->exec) {
    ENSCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(e);
    hipGraph_t g = nullptr;
    hipError_t he = hipStreamEndCapture(e->stream, &g);
    if (rc) return rc;
    if (he != hipSuccess) return solar::set_error(SOLAR_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(he));
    e->graph = g;
    ENSCHK(hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0));
  }
  for (int s = 0; s < nsteps; ++s) ENSCHK(hipGraphLaunch(e->exec, e->stream));
  return SOLAR_OK;
}

int solar_ens_sync(solar_ens* e) {
  if (!e) return solar::set_error(SOLAR_ERR_ARG, "solar_ens_sync: null handle");
  ENSCHK(hipStreamSynchronize(e->stream));
  return SOLAR_OK;
}

int solar_ens_rk4(solar_ens* e, int nsteps) {
  int rc = solar_ens_rk4_async(e, nsteps);
  return rc ? rc : solar_ens_sync(e);
}

int solar_ens_info(solar_ens* e, size_t* points, int* launches_per_step, size_t* device_bytes) {
  if (!e) return solar::set_error(SOLAR_ERR_ARG, "solar_ens_info: null handle");
  size_t p = 0;
  for (int m = 0; m < e->nm; ++m) p += static_cast<size_t>(6) * e->n[m] * e->n[m] * e->nr[m];
  if (points) *points = p;
  if (launches_per_step)
    *launches_per_step = 6 * static_cast<int>(e->gg.size()) + 5 * static_cast<int>(e->stage_l.size()) +
                         static_cast<int>(e->filter_l.size());
  if (device_bytes) *device_bytes = e->bytes;
  return SOLAR_OK;
}

}  // extern "C"
