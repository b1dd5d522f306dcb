// pn_batch.cpp -- many independent disks advanced together on one GPU.
//
// Each disk is an ordinary FP32 context (its own state, tables, dt, planet and scalars); the
// batch adds one stream and, per step, four launches for all of them: one workgroup per disk
// for the pre-step (Index_Shift, half kick, tables), one persistent launch per sweep over
// every disk's work items (radial: window x column group; azimuthal: workgroup of waves),
// and one workgroup per disk for the post-step (roll, planet, time, next dt from the CFL
// partials the azimuthal sweep wrote).  A step is captured as a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "penguin.h"
#include "pn_ctx.hpp"

using namespace pn;
using namespace pn::rt;

struct pn_batch {
  std::vector<pn_ctx*> disks;
  std::string err;
  hipStream_t stream = nullptr;
  int var_x = 0, var_y = 3;   // indices into kF32VariantsX / kF32VariantsY
  bool tuned = false;
  int grid_x = 0, grid_y = 0;
  DiskDev* dd = nullptr;      // device, one record per disk
  DevF32* dfs = nullptr;      // device
  int* order_x = nullptr;     // device, (disk, item) per slot of the radial launch (build_order_x)
  size_t order_cap = 0;       // ints order_x holds
  int nslot_x = 0;
  int* base_y = nullptr;     // device, ndisk + 1 azimuthal item offsets
  hipGraphExec_t gexec = nullptr;
  int gexec_cur = 0;          // the buffer the captured step reads (it returns there)
};

namespace {

int bfail(pn_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (b) b->err = buf;
  return code;
}

#define PB_HIP(b, call)                                                                         \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) return bfail((b), PN_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

int ndisk(const pn_batch* b) { return static_cast<int>(b->disks.size()); }

void drop_graph(pn_batch* b) {
  if (b->gexec) {
    hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
}

// a disk whose next dt cannot come from its last sweep: a moving planet or damping zones
bool cfl_before_step(const pn_ctx* c) { return c->P.plnt_flag == 2 || c->P.kill_flag != 0; }

// the batch's disks run in the barycentric frame (all or none): the kernels with its star terms
bool any_bary(const pn_batch* b) { return b->disks[0]->P.bary_flag != 0; }
// the batch's disks carry EOS 3 (all or none): the kernels with the fourth field
bool any_adia(const pn_batch* b) { return b->disks[0]->P.eos == 3; }

// The radial launch's items in the order its workgroups take them.  The persistent kernel's
// workgroup b runs slots b, b + G, ... (G a multiple of 8) on XCD b % 8, so slot t runs on
// XCD t % 8.  Each XCD gets a contiguous run of the disks' items (neighbouring column groups
// share L2 lines) of an equal share of the estimated cost, laid out by decreasing cost (a
// stable sort, so equal items keep their neighbours): every XCD's last, partial round then
// holds its cheapest items -- last windows short enough to run two columns per wavefront pass
// -- instead of a random mix, whose most expensive workgroup would set the launch's length.
// Cost of a window item: 0.3 for the staging (scaled by the window's cells) and 0.7 for the
// pencil passes (halved when the window runs two columns per pass, a third for three).
int build_order_x(pn_batch* b, std::vector<int>& order, int& nslots) {
  const int K = kF32VariantsX[b->var_x][0];
  const int nc = kF32VariantsX[b->var_x][1] * kF32VariantsX[b->var_x][2];
  const int real = 64 * K - 2 * kNpad;
  struct Item { int d, local; double cost; };
  std::vector<Item> items;
  double total = 0.0;
  for (int d = 0; d < ndisk(b); ++d) {
    const pn_ctx* c = b->disks[d];
    const int nwin = f32_windows(c->g.imax, K), ngrp = (c->g.jmax + nc - 1) / nc;
    for (int l = 0; l < nwin; ++l) {
      const int lim = (l < nwin - 1) ? 64 * K : c->g.imax - real * (nwin - 1) + 2 * kNpad;
      double passes = 1.0;
      if (c->df.pack_tails && K == 4 && lim <= 21 * K) passes = 1.0 / 3.0;
      else if (c->df.pack_tails && lim <= 32 * K) passes = 0.5;
      const double cost = 0.3 * lim / (64.0 * K) + 0.7 * passes;
      for (int g = 0; g < ngrp; ++g) {
        items.push_back({d, l * ngrp + g, cost});
        total += cost;
      }
    }
  }
  std::vector<std::vector<Item>> share(8);
  double cum = 0.0;
  for (const Item& it : items) {
    const int x = total > 0.0 ? std::min(7, static_cast<int>(8.0 * (cum + 0.5 * it.cost) / total)) : 0;
    share[x].push_back(it);
    cum += it.cost;
  }
  size_t len = 0;
  for (auto& s : share) {
    std::stable_sort(s.begin(), s.end(), [](const Item& a, const Item& c) { return a.cost > c.cost; });
    len = std::max(len, s.size());
  }
  nslots = static_cast<int>(8 * len);
  order.assign(2 * static_cast<size_t>(nslots), -1);
  for (int x = 0; x < 8; ++x)
    for (size_t i = 0; i < share[x].size(); ++i) {
      const size_t t = x + 8 * i;
      order[2 * t] = share[x][i].d;
      order[2 * t + 1] = share[x][i].local;
    }
  return PN_OK;
}

// the per-disk records and item offsets for the current variants
int upload_desc(pn_batch* b) {
  const int n = ndisk(b);
  std::vector<DiskDev> dd(n);
  std::vector<DevF32> df(n);
  std::vector<int> by(n + 1, 0);
  for (int d = 0; d < n; ++d) {
    pn_ctx* c = b->disks[d];
    if (!c->loaded) return bfail(b, PN_ERR_STATE, "disk %d has no state loaded", d);
    const int ny = f32_cfl_partials_y(c->dg, b->var_y);
    if (ny > c->ncflp) return bfail(b, PN_ERR_ARG, "disk %d: %d CFL partials exceed %d", d, ny, c->ncflp);
    DiskDev& D = dd[d];
    D.g = c->dg;
    D.p = c->dp;
    D.ss = c->ss;
    D.cflp = c->cflp;
    // a planet that moves changes FrRot inside the step, and damping zones change the state
    // after the sweeps: such a disk's CFL pass runs before the step (the next step's dt cannot
    // come from the last sweep)
    const bool pre = cfl_before_step(c);
    D.ncfl_post = pre ? 0 : ny;
    D.ncfl_pre = pre ? f32_cfl_blocks(c->dg) : 0;
    if (D.ncfl_pre > c->ncflp) return bfail(b, PN_ERR_ARG, "disk %d: %d CFL partials exceed %d", d, D.ncfl_pre, c->ncflp);
    D.trig = c->df.trig;
    D.tab = c->df.tab;
    D.tstride = c->df.tstride;
    df[d] = c->df;
    df[d].nwin_x = f32_windows(c->g.imax, kF32VariantsX[b->var_x][0]);
    df[d].nwin_y = f32_windows(c->g.jmax, kF32VariantsY[b->var_y][0]);
    if (c->df.tstride < c->dg.sr * c->dg.sk + 64 * kF32VariantsX[b->var_x][0])
      return bfail(b, PN_ERR_ARG, "disk %d: table overhang too short", d);
    by[d + 1] = by[d] + ny;
  }
  PB_HIP(b, hipMemcpyAsync(b->dd, dd.data(), n * sizeof(DiskDev), hipMemcpyHostToDevice, b->stream));
  PB_HIP(b, hipMemcpyAsync(b->dfs, df.data(), n * sizeof(DevF32), hipMemcpyHostToDevice, b->stream));
  PB_HIP(b, hipMemcpyAsync(b->base_y, by.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice, b->stream));
  std::vector<int> order;
  build_order_x(b, order, b->nslot_x);
  if (order.size() > b->order_cap) {   // a captured step names the old array: recapture
    drop_graph(b);
    PB_HIP(b, hipStreamSynchronize(b->stream));
    if (b->order_x) PB_HIP(b, hipFree(b->order_x));
    b->order_x = nullptr;
    b->order_cap = 0;
    PB_HIP(b, hipMalloc(reinterpret_cast<void**>(&b->order_x), order.size() * sizeof(int)));
    b->order_cap = order.size();
  }
  PB_HIP(b, hipMemcpyAsync(b->order_x, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
  PB_HIP(b, hipStreamSynchronize(b->stream));
  // persistent grids: what is resident at once, or fewer when there are fewer items
  auto fit = [](int resident, int items) { return std::max(8, std::min(resident, (items + 7) / 8 * 8)); };
  b->grid_x = fit(f32_batch_grid(0, b->var_x, any_bary(b), any_adia(b)), b->nslot_x);
  b->grid_y = fit(f32_batch_grid(1, b->var_y, any_bary(b), any_adia(b)), by[n]);
  if (b->grid_x < 8 || b->grid_y < 8) return bfail(b, PN_ERR_HIP, "occupancy query failed");
  return PN_OK;
}

// a disk's launch context on the batch's stream
LaunchCtx on_batch(pn_batch* b, pn_ctx* c) {
  LaunchCtx L = lctx(c);
  L.stream = b->stream;
  return L;
}

// one step of every disk: the CFL pass of each disk whose planet moves or which has damping
// zones, the batched pre-step, each disk's radial ghost copies, the batched sweeps, each
// damped disk's damping, the batched post-step,
// then each moving planet's time, drift, disk forces (modes 8 / 9 from its FP32 state) and
// closing kick -- the order pn_run gives a single disk (damping is ring-uniform, so it commutes
// with the roll that the post-step does for a moving planet's disk)
int launch_batch_step(pn_batch* b, int cur) {
  const int n = ndisk(b);
  hipError_t e = hipSuccess;
  for (pn_ctx* c : b->disks) {
    if (!cfl_before_step(c) || e != hipSuccess) continue;
    int ncfl = 0;
    e = f32_cfl(on_batch(b, c), c->df, cur, &ncfl);
  }
  if (e == hipSuccess) e = launch_pre_step_batch(b->stream, b->dd, n);
  for (pn_ctx* c : b->disks) {   // radial boundary codes other than 3 (launches nothing for 3)
    if (e != hipSuccess) break;
    e = f32_fill_ghosts(on_batch(b, c), c->df, cur);
  }
  if (e == hipSuccess) e = f32_sweep_x_batch(b->stream, b->dd, b->dfs, b->order_x, b->nslot_x, b->var_x, cur, b->grid_x, any_bary(b), any_adia(b));
  if (e == hipSuccess) e = f32_sweep_y_batch(b->stream, b->dd, b->dfs, b->base_y, n, b->var_y, cur ^ 1, b->grid_y, any_bary(b), any_adia(b));
  for (pn_ctx* c : b->disks) {
    if (!c->P.kill_flag || e != hipSuccess) continue;
    e = f32_damp(on_batch(b, c), c->df, cur);
  }
  if (e == hipSuccess) e = launch_post_step_batch(b->stream, b->dd, n);
  for (pn_ctx* c : b->disks) {
    if (c->P.plnt_flag != 2 || e != hipSuccess) continue;
    e = launch_planet_after_sweep(on_batch(b, c), cur, &c->df);
  }
  if (e != hipSuccess) return bfail(b, PN_ERR_HIP, "batched step: %s", hipGetErrorString(e));
  return PN_OK;
}

// n steps of every disk from buffer cur (a 2-D step returns there): replays of the step
// captured as a graph from that buffer, or direct launches
int batch_steps(pn_batch* b, int cur, int nsteps, int graph) {
  int done = 0, rc;
  if (graph && nsteps > 0) {
    // the captured step names its buffer; a disk moved to the other one since (for example by
    // pn_sweep_axis on pn_batch_disk) needs a step captured from there
    if (b->gexec && b->gexec_cur != cur) drop_graph(b);
    if (!b->gexec) {
      hipGraph_t gr;
      PB_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      rc = launch_batch_step(b, cur);
      hipError_t e = hipStreamEndCapture(b->stream, &gr);
      if (rc != PN_OK) return rc;
      if (e != hipSuccess) return bfail(b, PN_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&b->gexec, gr, nullptr, nullptr, 0);
      hipGraphDestroy(gr);
      if (e != hipSuccess) return bfail(b, PN_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
      b->gexec_cur = cur;
    }
    for (; done < nsteps; ++done) PB_HIP(b, hipGraphLaunch(b->gexec, b->stream));
  }
  for (; done < nsteps; ++done)
    if ((rc = launch_batch_step(b, cur)) != PN_OK) return rc;
  return PN_OK;
}

int common_cur(pn_batch* b, int* cur) {
  *cur = b->disks[0]->cur;
  for (pn_ctx* c : b->disks) {
    if (!c->loaded) return bfail(b, PN_ERR_STATE, "a disk has no state loaded");
    if (c->cur != *cur) return bfail(b, PN_ERR_STATE, "disks sit in different buffers");
  }
  return PN_OK;
}

// per-step tables of every disk (trig, vrot) without touching the state
int prep_all(pn_batch* b) {
  for (pn_ctx* c : b->disks) {
    LaunchCtx L = lctx(c);
    L.stream = b->stream;
    hipError_t e = f32_prep(L, c->df);
    if (e != hipSuccess) return bfail(b, PN_ERR_HIP, "prep: %s", hipGetErrorString(e));
  }
  return PN_OK;
}

template <class F>
int btime(pn_batch* b, int nrep, double* ms, F&& launch) {
  hipEvent_t e0, e1;
  PB_HIP(b, hipEventCreate(&e0));
  PB_HIP(b, hipEventCreate(&e1));
  hipError_t e = hipSuccess;
  for (int k = 0; k < nrep && e == hipSuccess; ++k) e = launch();   // untimed: clocks back to speed
  if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
  for (int k = 0; k < nrep && e == hipSuccess; ++k) e = launch();
  if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (e != hipSuccess) return bfail(b, PN_ERR_HIP, "batch timing: %s", hipGetErrorString(e));
  *ms = static_cast<double>(t) / nrep;
  return PN_OK;
}

bool any_viscous(const pn_batch* b) {
  for (const pn_ctx* c : b->disks)
    if (c->P.visc_flag) return true;
  return false;
}

}  // namespace

extern "C" {

int pn_batch_create(int n, const pn_grid* grids, const pn_params* params, int precision, pn_batch** out) {
  if (!out) return PN_ERR_ARG;
  *out = nullptr;
  pn_batch* b = new pn_batch();
  *out = b;   // the caller reads the message of a failure, then pn_batch_destroy
  if (n < 1 || !grids || !params) return bfail(b, PN_ERR_ARG, "batch of %d disks", n);
  if (precision != PN_FP32) return bfail(b, PN_ERR_UNSUPPORTED, "the batched launch is built for the FP32 path");
  // the barycentric frame's star terms are compiled into kernels of their own; one launch runs
  // one kind, so the disks of a batch share the frame (a compile-time choice of the reference,
  // forces.cu:28), and the closure (global.h:18)
  for (int d = 1; d < n; ++d) {
    if ((params[d].bary_flag != 0) != (params[0].bary_flag != 0))
      return bfail(b, PN_ERR_UNSUPPORTED, "disk %d: the disks of a batch share the frame (bary_flag)", d);
    // likewise the closure: EOS 3 disks carry a fourth field through kernels of their own
    if ((params[d].eos == 3) != (params[0].eos == 3))
      return bfail(b, PN_ERR_UNSUPPORTED, "disk %d: the disks of a batch share the closure (eos)", d);
  }
  for (int d = 0; d < n; ++d) {
    if (params[d].ndim != 2)
      return bfail(b, PN_ERR_UNSUPPORTED, "disk %d: batches are built for 2-D disks", d);
    if (params[d].eos != 0 && params[d].eos != 3)
      return bfail(b, PN_ERR_UNSUPPORTED, "disk %d: batches are built for EOS 0 and EOS 3", d);
    pn_ctx* c = nullptr;
    const int rc = pn_create(&grids[d], &params[d], precision, &c);
    if (c) b->disks.push_back(c);
    if (rc != PN_OK) return bfail(b, rc, "disk %d: %s", d, c ? c->err.c_str() : "pn_create failed");
  }
  if (any_viscous(b)) b->var_y = 1;   // (2, 4): the reference's azimuthal windows
  PB_HIP(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  PB_HIP(b, hipMalloc(reinterpret_cast<void**>(&b->dd), n * sizeof(DiskDev)));
  PB_HIP(b, hipMalloc(reinterpret_cast<void**>(&b->dfs), n * sizeof(DevF32)));
  PB_HIP(b, hipMalloc(reinterpret_cast<void**>(&b->base_y), (n + 1) * sizeof(int)));
  return PN_OK;
}

int pn_batch_destroy(pn_batch* b) {
  if (!b) return PN_OK;
  if (b->stream) hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (pn_ctx* c : b->disks) pn_destroy(c);
  if (b->dd) hipFree(b->dd);
  if (b->dfs) hipFree(b->dfs);
  if (b->order_x) hipFree(b->order_x);
  if (b->base_y) hipFree(b->base_y);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return PN_OK;
}

const char* pn_batch_last_error(const pn_batch* b) { return b ? b->err.c_str() : "null batch"; }

pn_ctx* pn_batch_disk(pn_batch* b, int d) {
  if (!b || d < 0 || d >= ndisk(b)) return nullptr;
  return b->disks[d];
}

int pn_batch_size(const pn_batch* b) { return b ? ndisk(b) : 0; }

int pn_batch_run_begin(pn_batch* b, pn_run_state* states) {
  if (!b || !states) return PN_ERR_ARG;
  for (int d = 0; d < ndisk(b); ++d) {
    const int rc = pn_run_begin(b->disks[d], &states[d]);
    if (rc != PN_OK) return bfail(b, rc, "disk %d: %s", d, b->disks[d]->err.c_str());
  }
  return PN_OK;
}

int pn_batch_run_restart(pn_batch* b, int startat, pn_run_state* states) {
  if (!b || !states) return PN_ERR_ARG;
  for (int d = 0; d < ndisk(b); ++d) {
    const int rc = pn_run_restart(b->disks[d], startat, &states[d]);
    if (rc != PN_OK) return bfail(b, rc, "disk %d: %s", d, b->disks[d]->err.c_str());
  }
  return PN_OK;
}

int pn_batch_set_variants(pn_batch* b, int var_x, int var_y) {
  if (!b) return PN_ERR_ARG;
  if (var_x < 0 || var_x >= kNumF32Variants || var_y < 0 || var_y >= kNumF32Variants)
    return bfail(b, PN_ERR_ARG, "variants (%d, %d) outside [0, %d)", var_x, var_y, kNumF32Variants);
  if (!f32_y_variant_allowed(var_y, any_viscous(b)))
    return bfail(b, PN_ERR_ARG, "azimuthal variant %d: viscous disks keep 128-cell azimuthal windows", var_y);
  if (var_x != b->var_x || var_y != b->var_y) drop_graph(b);
  b->var_x = var_x;
  b->var_y = var_y;
  b->tuned = true;
  return PN_OK;
}

int pn_batch_get_variants(const pn_batch* b, int* var_x, int* var_y) {
  if (!b || !var_x || !var_y) return PN_ERR_ARG;
  *var_x = b->var_x;
  *var_y = b->var_y;
  return PN_OK;
}

int pn_batch_tune(pn_batch* b, int nrep, int* var_x, int* var_y) {
  if (!b || nrep < 1) return PN_ERR_ARG;
  int cur, rc;
  if ((rc = common_cur(b, &cur)) != PN_OK) return rc;
  if ((rc = prep_all(b)) != PN_OK) return rc;
  const int keep_x = b->var_x, keep_y = b->var_y;
  double best_x = 1e30, best_y = 1e30;
  int bx = keep_x, by = keep_y;
  // every candidate reads the current buffer and writes the other one: the state stays
  for (int v = 0; v < kNumF32Variants; ++v) {
    b->var_x = v;
    b->var_y = v;
    if ((rc = upload_desc(b)) != PN_OK) return rc;
    double tx = 0.0, ty = 0.0;
    const int n = ndisk(b);
    rc = btime(b, nrep, &tx, [&]() {
      return f32_sweep_x_batch(b->stream, b->dd, b->dfs, b->order_x, b->nslot_x, b->var_x, cur, b->grid_x, any_bary(b), any_adia(b));
    });
    const bool y_ok = f32_y_variant_allowed(v, any_viscous(b));
    if (rc == PN_OK && y_ok)
      rc = btime(b, nrep, &ty, [&]() {
        return f32_sweep_y_batch(b->stream, b->dd, b->dfs, b->base_y, n, b->var_y, cur, b->grid_y, any_bary(b), any_adia(b));
      });
    if (rc != PN_OK) return rc;
    if (tx < best_x) { best_x = tx; bx = v; }
    if (y_ok && ty < best_y) { best_y = ty; by = v; }
  }
  b->var_x = bx;
  b->var_y = by;
  b->tuned = true;
  drop_graph(b);
  if (var_x) *var_x = bx;
  if (var_y) *var_y = by;
  return PN_OK;
}

int pn_batch_run(pn_batch* b, int nsteps, int graph, pn_run_state* states) {
  if (!b || !states || nsteps < 0) return PN_ERR_ARG;
  int cur, rc;
  if ((rc = common_cur(b, &cur)) != PN_OK) return rc;
  const int n = ndisk(b);
  bool land = false;
  for (int d = 0; d < n; ++d) {
    if (!states[d].land) continue;
    land = true;
    // as pn_run: a landing needs time left before both marks
    const pn_params& P = b->disks[d]->P;
    if (!(states[d].timem < P.tmovie && states[d].simtime < P.endtime))
      return bfail(b, PN_ERR_ARG, "disk %d: landing with no time left: timem %g of tmovie %g, simtime %g of endtime %g",
                   d, states[d].timem, P.tmovie, states[d].simtime, P.endtime);
  }
  if (!b->tuned && (rc = pn_batch_tune(b, 3, nullptr, nullptr)) != PN_OK) return rc;
  if ((rc = upload_desc(b)) != PN_OK) return rc;
  // every disk's scalars; a disk whose CFL is fused into its last sweep takes the first dt
  // from its state here (the sweeps supply the rest), one whose planet moves in its pre-step
  std::vector<StepScalars> h(n);
  for (int d = 0; d < n; ++d) {
    pn_ctx* c = b->disks[d];
    h[d].dt = states[d].dt;
    h[d].FrRot = states[d].FrRot;
    h[d].simtime = states[d].simtime;
    h[d].planet = states[d].planet;
    h[d].ncycle = states[d].ncycle;
    h[d].timem = states[d].timem;
    h[d].land = states[d].land ? 1 : 0;
    h[d].event = 0;
    if ((rc = upload_scalars(c, h[d])) != PN_OK ||
        (nsteps > 0 && fused_cfl(c) && (rc = launch_cfl_any(c, cur)) != PN_OK) || (rc = sync(c)) != PN_OK)
      return bfail(b, rc, "disk %d: %s", d, c->err.c_str());
  }
  auto download_all = [&]() -> int {
    for (int d = 0; d < n; ++d) {
      pn_ctx* c = b->disks[d];
      const int r = download_scalars(c, &h[d]);
      if (r != PN_OK) return bfail(b, r, "disk %d: %s", d, c->err.c_str());
    }
    return PN_OK;
  };
  if (!land) {
    if ((rc = batch_steps(b, cur, nsteps, graph)) != PN_OK) return rc;
  } else {
    // the landing is applied on the device, per disk, every step; the host stops after a step
    // in which any disk landed, enqueuing at once only the steps that cannot reach one
    int done = 0;
    while (done < nsteps) {
      int m = nsteps - done;
      for (int d = 0; d < n; ++d) {
        if (!h[d].land) continue;
        const pn_params& P = b->disks[d]->P;
        const double room = std::fmin(P.tmovie - h[d].timem, P.endtime - h[d].simtime);
        m = std::min(m, std::max(1, steps_before_landing(h[d].dt, room, nsteps - done)));
      }
      if ((rc = batch_steps(b, cur, m, graph)) != PN_OK) return rc;
      done += m;
      if ((rc = download_all()) != PN_OK) return rc;
      bool stop = false;
      for (int d = 0; d < n; ++d) stop = stop || h[d].event != 0 || h[d].bad;
      if (stop) break;
    }
  }
  PB_HIP(b, hipStreamSynchronize(b->stream));
  if ((rc = download_all()) != PN_OK) return rc;
  for (int d = 0; d < n; ++d) {
    // a fused CFL leaves the next step's dt in the record: report the last one used
    states[d].dt = (fused_cfl(b->disks[d]) && nsteps > 0) ? h[d].dt_prev : h[d].dt;
    states[d].FrRot = h[d].FrRot;
    states[d].simtime = h[d].simtime;
    states[d].planet = h[d].planet;
    states[d].ncycle = h[d].ncycle;
    states[d].timem = h[d].timem;
    states[d].event = h[d].event;
    if (h[d].bad) return bfail(b, PN_ERR_NUMERIC, "disk %d: non-finite or non-positive dt during the run", d);
  }
  return PN_OK;
}

int pn_batch_time_kernels(pn_batch* b, int nrep, double* ms4) {
  if (!b || !ms4 || nrep < 1) return PN_ERR_ARG;
  int cur, rc;
  if ((rc = common_cur(b, &cur)) != PN_OK) return rc;
  if ((rc = upload_desc(b)) != PN_OK) return rc;
  const int n = ndisk(b);
  rc = btime(b, nrep, &ms4[0], [&]() { return launch_pre_step_batch(b->stream, b->dd, n); });
  if (rc == PN_OK)
    rc = btime(b, nrep, &ms4[1], [&]() {
      return f32_sweep_x_batch(b->stream, b->dd, b->dfs, b->order_x, b->nslot_x, b->var_x, cur, b->grid_x, any_bary(b), any_adia(b));
    });
  if (rc == PN_OK)
    rc = btime(b, nrep, &ms4[2], [&]() {
      return f32_sweep_y_batch(b->stream, b->dd, b->dfs, b->base_y, n, b->var_y, cur, b->grid_y, any_bary(b), any_adia(b));
    });
  if (rc == PN_OK) rc = btime(b, nrep, &ms4[3], [&]() { return launch_post_step_batch(b->stream, b->dd, n); });
  return rc;
}

}  // extern "C"
