// pn_runtime.cpp -- the C ABI: contexts, hydr_ring upload/download, the step operators and
// the device-resident step loop (optionally replayed as a HIP graph), for the plain FP64
// path and the bulk-FP32 path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "penguin.h"
#include "pn_f32.hpp"
#include "pn_host.hpp"
#include "pn_launch.hpp"
#include "pn_ctx.hpp"

using namespace pn;
using namespace pn::rt;

namespace {
constexpr double kPiD = 3.1415926535897932384626433832795;
}


namespace pn {
namespace rt {

int fail(pn_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define PN_HIP(c, call)                                                                        \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) return fail((c), PN_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define PN_LAUNCH(c, what, call)                                                          \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) return fail((c), PN_ERR_HIP, "%s: %s", what, hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int dalloc(pn_ctx* c, T** p, size_t n) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, n * sizeof(T) + 64);
  if (e != hipSuccess) return fail(c, PN_ERR_HIP, "hipMalloc(%zu): %s", n * sizeof(T), hipGetErrorString(e));
  c->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return PN_OK;
}

LaunchCtx lctx(pn_ctx* c) {
  LaunchCtx L;
  L.g = c->dg;
  L.p = c->dp;
  L.ss = c->ss;
  L.partials = c->partials;
  L.cflp = c->cflp;
  L.ncflp = c->ncflp;
  L.stream = c->stream;
  return L;
}

int nwin(int len) { return (len + kReal - 1) / kReal; }

bool is_f32(const pn_ctx* c) { return c->precision == PN_FP32; }

// the next step's CFL maximum can be taken in the last sweep's epilogue when FrRot cannot
// change inside a step, i.e. unless the planet moves under the disk's torque
// the next dt from the last sweep's epilogue: FP32, no work after the sweeps that changes the
// state the CFL reads (a moving planet's update reads the sweeps' result; damping changes it)
bool fused_cfl(const pn_ctx* c) { return is_f32(c) && c->P.plnt_flag < 2 && !c->P.kill_flag; }

// the FP32 radial sweep of the context's geometry
hipError_t f32_x(const pn_ctx* c, const LaunchCtx& L, int variant, int src) {
  return (c->P.ndim == 3) ? f32_sweep_x3(L, c->df, variant, src) : f32_sweep_x(L, c->df, variant, src);
}

// CFL partials the step's last FP32 sweep writes
int f32_fused_partials(const pn_ctx* c) {
  return (c->P.ndim == 3) ? f32_cfl_partials_z(c->dg, c->var_z) : f32_cfl_partials_y(c->dg, c->var_y);
}

int check_params(pn_ctx* c, const pn_grid* g, const pn_params* p, int precision) {
  if (g->imax < 1 || g->jmax < 2 * kNpad || g->kmax < 1) return fail(c, PN_ERR_ARG, "bad grid %d x %d x %d", g->imax, g->jmax, g->kmax);
  if (p->ndim != 2 && p->ndim != 3) return fail(c, PN_ERR_ARG, "ndim must be 2 or 3");
  if ((p->ndim == 2) != (g->kmax == 1)) return fail(c, PN_ERR_ARG, "ndim %d with kmax %d", p->ndim, g->kmax);
  if (p->eos < 0 || p->eos > 3) return fail(c, PN_ERR_ARG, "eos %d", p->eos);
  const bool cyl = (p->ngeomx == 1 && p->ngeomy == 3 && p->ngeomz == 0);
  const bool sph = (p->ngeomx == 2 && p->ngeomy == 4 && p->ngeomz == 5);
  if (p->ndim == 2 && !cyl) return fail(c, PN_ERR_UNSUPPORTED, "2-D needs cylindrical geometry (1,3,0)");
  if (p->ndim == 3 && !sph) return fail(c, PN_ERR_UNSUPPORTED, "3-D needs spherical geometry (2,4,5)");
  if (p->ndim == 3 && (p->eos == 1 || p->eos == 3)) return fail(c, PN_ERR_UNSUPPORTED, "3-D EOS %d is not built", p->eos);
  if (p->fargo_flag < 0 || p->fargo_flag > 2) return fail(c, PN_ERR_ARG, "fargo_flag %d", p->fargo_flag);
  if (p->plnt_flag < 0 || p->plnt_flag > 2) return fail(c, PN_ERR_ARG, "plnt_flag %d", p->plnt_flag);
  // a moving planet in 3-D: the disk-on-planet reductions (modes 8 / 9) are 2-D only in the
  // reference (output_lv1, device_func.cu:81-108, returns 0 for them in 3-D), so the planet
  // moves under the star alone, as there
  if (p->nlft < 0 || p->nlft > 4 || p->nrgh < 0 || p->nrgh > 4)
    return fail(c, PN_ERR_ARG, "radial boundary codes (%d, %d) outside 0..4", p->nlft, p->nrgh);
  if (p->kill_flag != 0 && p->kill_flag != 1) return fail(c, PN_ERR_ARG, "kill_flag %d", p->kill_flag);
  for (int code : {p->nlft, p->nrgh})
    if ((code == 0 || code == 2) && g->imax < kNpad + 1)
      return fail(c, PN_ERR_ARG, "radial boundary code %d needs imax > %d", code, kNpad);
  if (precision == PN_FP32 && p->eos != 0 && !(p->eos == 3 && p->ndim == 2))
    return fail(c, PN_ERR_UNSUPPORTED, "the FP32 path is built for EOS 0 and the 2-D EOS 3 disk");
  if (p->bary_flag != 0 && p->bary_flag != 1) return fail(c, PN_ERR_ARG, "bary_flag %d", p->bary_flag);
  return PN_OK;
}

int upload_scalars(pn_ctx* c, const StepScalars& h) {   // synchronous: h is a caller's stack record
  PN_HIP(c, hipMemcpyAsync(c->ss, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
  PN_HIP(c, hipStreamSynchronize(c->stream));
  return PN_OK;
}

int download_scalars(pn_ctx* c, StepScalars* h) {
  PN_HIP(c, hipMemcpyAsync(h, c->ss, sizeof *h, hipMemcpyDeviceToHost, c->stream));
  PN_HIP(c, hipStreamSynchronize(c->stream));
  return PN_OK;
}

int sync(pn_ctx* c) {
  PN_HIP(c, hipStreamSynchronize(c->stream));
  return PN_OK;
}

int launch_cfl_any(pn_ctx* c, int src) {   // partial maxima, then dt
  LaunchCtx L = lctx(c);
  if (is_f32(c)) {
    int n = 0;
    PN_LAUNCH(c, "cfl", f32_cfl(L, c->df, src, &n));
    PN_LAUNCH(c, "dt", launch_dt_finalize(L, n));
  } else {
    PN_LAUNCH(c, "cfl", launch_cfl(L, src));
  }
  return PN_OK;
}

// sweeps of one step; the state returns to cur in 2-D and moves to the other buffer in 3-D.
// prep: refresh the FP32 per-step tables first; roll: the integer roll after the sweeps.
int launch_sweeps(pn_ctx* c, int& cur, int do_cfl, bool prep, bool roll) {
  LaunchCtx L = lctx(c);
  if (is_f32(c)) {
    if (prep) PN_LAUNCH(c, "prep", f32_prep(L, c->df));
    PN_LAUNCH(c, "ghosts", f32_fill_ghosts(L, c->df, cur));
    PN_LAUNCH(c, "sweep_x", f32_x(c, L, c->var_x, cur));
    if (c->P.ndim == 3) {
      PN_LAUNCH(c, "sweep_y", f32_sweep_y(L, c->df, c->var_y, cur ^ 1, 0));
      PN_LAUNCH(c, "sweep_z", f32_sweep_z(L, c->df, c->var_z, cur, do_cfl));
      cur ^= 1;
    } else {
      PN_LAUNCH(c, "sweep_y", f32_sweep_y(L, c->df, c->var_y, cur ^ 1, do_cfl));
    }
  } else {
    PN_LAUNCH(c, "ghosts", launch_fill_ghosts(L, cur));
    PN_LAUNCH(c, "sweep_x", fp64_sweep_x(L, cur));
    PN_LAUNCH(c, "sweep_y", fp64_sweep_y(L, cur ^ 1));
    if (c->P.ndim == 3) {
      PN_LAUNCH(c, "sweep_z", fp64_sweep_z(L, cur));
      cur ^= 1;
    }
  }
  if (roll && c->P.fargo_flag > 0) PN_LAUNCH(c, "rotate", launch_rotate(L));
  return PN_OK;
}

// one full step of main.cu:345-387 on the device, reading and writing c->ss:
//   [CFL partials] pre-step (dt, Index_Shift, half kick, FP32 tables) -> sweeps -> post-step
// (roll, planet, time, and the next dt when the CFL is fused into the last sweep)
int launch_step(pn_ctx* c, int& cur) {
  LaunchCtx L = lctx(c);
  const bool fused = fused_cfl(c);
  int ncfl = 0, rc;
  if (!fused) {
    if (is_f32(c)) PN_LAUNCH(c, "cfl", f32_cfl(L, c->df, cur, &ncfl));
    else PN_LAUNCH(c, "cfl", launch_cfl_partials(L, cur, &ncfl));
  }
  float2* trig = is_f32(c) ? c->df.trig : nullptr;
  float* tab = is_f32(c) ? c->df.tab : nullptr;
  float* tabz = is_f32(c) ? c->df.tabz : nullptr;
  PN_LAUNCH(c, "pre step", launch_pre_step(L, ncfl, trig, tab, tabz, c->df.tstride));
  const bool post = c->P.plnt_flag < 2;
  if ((rc = launch_sweeps(c, cur, fused ? 1 : 0, false, !post)) != PN_OK) return rc;
  // damping zones: ring-uniform relaxation, which commutes with the integer roll
  if (c->P.kill_flag) PN_LAUNCH(c, "damp", is_f32(c) ? f32_damp(L, c->df, cur) : launch_damp(L, cur));
  if (post) PN_LAUNCH(c, "post step", launch_post_step(L, fused ? f32_fused_partials(c) : 0));
  else PN_LAUNCH(c, "planet update", launch_planet_after_sweep(L, cur, is_f32(c) ? &c->df : nullptr));
  return PN_OK;
}

}  // namespace rt
}  // namespace pn

extern "C" {

const char* pn_last_error(const pn_ctx* c) { return c ? c->err.c_str() : "null context"; }

int pn_create(const pn_grid* g, const pn_params* p, int precision, pn_ctx** out) {
  if (!g || !p || !out) return PN_ERR_ARG;
  *out = nullptr;
  pn_ctx* c = new pn_ctx();
  int rc = (precision != PN_FP64 && precision != PN_FP32) ? fail(c, PN_ERR_ARG, "precision %d", precision)
                                                           : check_params(c, g, p, precision);
  if (rc != PN_OK) {
    *out = c;   // the caller can read the message, then pn_destroy
    return rc;
  }
  c->g = *g;
  c->P = *p;
  c->precision = precision;
  if (p->visc_flag) c->var_y = 1;   // (2, 4): the reference's azimuthal windows
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    *out = c;
    return fail(c, PN_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  {
    // set-up of the no-allocation entry points: this device's scratch address, once
    double* scr = nullptr;
    size_t nscr = 0;
    e = noalloc_scratch(&scr, &nscr);
    if (e != hipSuccess) {
      *out = c;
      return fail(c, PN_ERR_HIP, "scratch address (hipGetSymbolAddress): %s", hipGetErrorString(e));
    }
  }
  DevGrid& d = c->dg;
  d.imax = g->imax;
  d.jmax = g->jmax;
  d.kmax = g->kmax;
  d.ndim = p->ndim;
  d.sr = g->imax + 2 * kNpad;
  d.sk = (p->ndim == 3) ? g->kmax + 2 * kNpad : 1;
  d.k0 = (p->ndim == 3) ? kNpad : 0;
  d.nwin_x = nwin(g->imax);
  d.nwin_y = nwin(g->jmax);
  d.nwin_z = nwin(g->kmax);
  const long nslot = static_cast<long>(d.sr) * d.sk;
  const size_t ncell = static_cast<size_t>(nslot) * g->jmax;
  // the FP32 sweeps address a field buffer with 32-bit byte offsets
  if (precision == PN_FP32 && ncell * sizeof(float) >= (size_t(1) << 31)) {
    *out = c;
    return fail(c, PN_ERR_UNSUPPORTED, "FP32 field buffers of %zu bytes: the FP32 path addresses below 2 GiB",
                ncell * sizeof(float));
  }
  int r2 = PN_OK;
  if (precision == PN_FP64) {
    for (int b = 0; b < 2 && r2 == PN_OK; ++b)
      for (int f = 0; f < kNField && r2 == PN_OK; ++f) r2 = dalloc(c, &d.f[b][f], ncell);
  } else {
    for (int b = 0; b < 2 && r2 == PN_OK; ++b)
      for (int f = 0; f < (p->ndim == 3 || p->eos == 3 ? 4 : 3) && r2 == PN_OK; ++f) r2 = dalloc(c, &c->df.f[b][f], ncell);
  }
  double *x, *dx, *xc, *xvol, *z, *dz, *zc, *zvol, *radc, *cs2, *nu, *y, *dy, *yc, *yvol;
  pn_symdisk* val;
  double** slot_tabs[] = {&x, &dx, &xc, &xvol, &z, &dz, &zc, &zvol, &radc, &cs2, &nu, &d.rot_v, &d.res_v};
  for (double** t : slot_tabs)
    if (r2 == PN_OK) r2 = dalloc(c, t, nslot);
  if (r2 == PN_OK) r2 = dalloc(c, &d.rot_j, nslot);
  if (r2 == PN_OK) r2 = dalloc(c, &d.inc_j, nslot);
  double** col_tabs[] = {&y, &dy, &yc, &yvol};
  for (double** t : col_tabs)
    if (r2 == PN_OK) r2 = dalloc(c, t, g->jmax);
  if (r2 == PN_OK) r2 = dalloc(c, &val, static_cast<size_t>(g->imax) * g->kmax);
  if (r2 == PN_OK) r2 = dalloc(c, &c->ss, 1);
  if (r2 == PN_OK) r2 = dalloc(c, &c->partials, kMaxPartials);
  if (r2 == PN_OK) r2 = dalloc(c, &c->out2, 2);
  if (r2 == PN_OK) r2 = dalloc(c, &c->ringbuf, static_cast<size_t>(g->imax) * g->kmax);
  // CFL partials: one per block of the grid-stride CFL kernels (at most 2048) or one per
  // workgroup of the FP32 sweep that takes the CFL in its epilogue, in any variant
  c->ncflp = 2048;
  if (precision == PN_FP32)
    for (int v = 0; v < kNumF32VariantsZ; ++v) {
      if (v < kNumF32Variants) c->ncflp = std::max(c->ncflp, f32_cfl_partials_y(d, v));
      if (p->ndim == 3) c->ncflp = std::max(c->ncflp, f32_cfl_partials_z(d, v));
    }
  if (r2 == PN_OK) r2 = dalloc(c, &c->cflp, c->ncflp);
  if (precision == PN_FP32) {
    float *fdy, *fyvol;
    double *rbg, *vbg;
    c->df.tstride = static_cast<int>((nslot + kTabOverhang + 3) / 4 * 4);
    if (r2 == PN_OK) r2 = dalloc(c, &c->df.tab, static_cast<size_t>(kTabEntries) * c->df.tstride);
    if (r2 == PN_OK && p->ndim == 3) r2 = dalloc(c, &c->df.tabz, static_cast<size_t>(kTabEntries) * c->df.tstride);
    if (r2 == PN_OK) r2 = dalloc(c, &rbg, nslot);
    if (r2 == PN_OK) r2 = dalloc(c, &vbg, nslot);
    if (r2 == PN_OK) r2 = dalloc(c, &fdy, g->jmax);
    if (r2 == PN_OK) r2 = dalloc(c, &fyvol, g->jmax);
    if (r2 == PN_OK) r2 = dalloc(c, &c->df.trig, g->jmax);
    if (r2 == PN_OK) {
      c->df.rbg = rbg;
      c->df.vbg = vbg;
      c->df.dy = fdy;
      c->df.yvol = fyvol;
    }
  }
  if (r2 != PN_OK) {
    *out = c;
    return r2;
  }
  d.x = x; d.dx = dx; d.xc = xc; d.xvol = xvol; d.z = z; d.dz = dz; d.zc = zc; d.zvol = zvol;
  d.radc = radc; d.cs2 = cs2; d.nu = nu; d.y = y; d.dy = dy; d.yc = yc; d.yvol = yvol; d.val = val;
  DevParams& q = c->dp;
  q.gam = p->gam;
  q.courant = p->courant;
  q.sc_h = p->sc_h;
  q.p_alpha = p->p_alpha;
  q.p_beta = p->p_beta;
  q.vis_nu = p->vis_nu;
  q.M_p = p->M_p;
  q.t_growth = p->t_growth;
  q.rs_fac = p->rs_fac;
  q.eos = p->eos;
  q.flat = p->flat_flag;
  q.plnt = p->plnt_flag;
  q.visc = p->visc_flag;
  q.fargo = p->fargo_flag;
  q.frrot = p->frrot_flag;
  q.ngeomx = p->ngeomx;
  q.ngeomy = p->ngeomy;
  q.ngeomz = p->ngeomz;
  q.ndim = p->ndim;
  q.has_w = (p->ndim == 3) ? 1 : 0;
  q.nlft = p->nlft;
  q.nrgh = p->nrgh;
  q.kill = p->kill_flag;
  q.bary = p->bary_flag;
  q.xmin = p->xmin;
  q.xmax = p->xmax;
  q.endtime = p->endtime;
  q.tmovie = p->tmovie;
  StepScalars s0{};
  rc = upload_scalars(c, s0);
  *out = c;
  return rc;
}

int pn_destroy(pn_ctx* c) {
  if (!c) return PN_OK;
  // drain first: no replay of the graph, and no kernel reading these buffers, may still run
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  for (void* q : c->allocs) hipFree(q);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return PN_OK;
}

int pn_load_rings(pn_ctx* c, const void* rings, size_t rings_bytes, const void* lft, const void* rgh, size_t lr_bytes,
                  const void* udr, const void* top, size_t ut_bytes, const pn_symdisk* val, size_t nval) {
  if (!c) return PN_ERR_ARG;
  if (c->ss == nullptr) return fail(c, PN_ERR_STATE, "context was not created");
  const int I = c->g.imax, J = c->g.jmax, K = c->g.kmax;
  const size_t rb = pn_ring_bytes(J);
  const size_t nring = static_cast<size_t>(I) * K;
  if (!rings || rings_bytes != rb * nring) return fail(c, PN_ERR_ARG, "rings: %zu bytes, expected %zu", rings_bytes, rb * nring);
  if (!lft || !rgh || lr_bytes != rb * PN_NPAD * K) return fail(c, PN_ERR_ARG, "lft/rgh: %zu bytes, expected %zu", lr_bytes, rb * PN_NPAD * K);
  if (c->P.ndim == 3 && (!udr || !top || ut_bytes != rb * PN_NPAD * I))
    return fail(c, PN_ERR_ARG, "udr/top: %zu bytes, expected %zu", ut_bytes, rb * PN_NPAD * I);
  if (!val || nval != nring) return fail(c, PN_ERR_ARG, "val: %zu records, expected %zu", nval, nring);

  const DevGrid& d = c->dg;
  const long nslot = static_cast<long>(d.sr) * d.sk;
  std::vector<double> F[kNField];
  for (int f = 0; f < kNField; ++f) F[f].assign(static_cast<size_t>(nslot) * J, 0.0);
  std::vector<double> sx(nslot, 0.0), sdx(nslot, 1.0), sxc(nslot, 1.0), sxvol(nslot, 1.0), sz(nslot, 0.0),
      sdz(nslot, 1.0), szc(nslot, 0.0), szvol(nslot, 1.0), srot_v(nslot, 0.0), sres_v(nslot, 0.0), svbg(nslot, 0.0);
  std::vector<int32_t> srot_j(nslot, 0), sinc_j(nslot, 0);
  // field values are copied as bytes so that signed zeros and NaN payloads survive
  auto take = [&](const void* base, long idx, int slot, bool interior) -> int {
    RingView R = ring_view(const_cast<void*>(base), J, idx);
    std::memcpy(&F[FR][static_cast<size_t>(slot) * J], R.r, 8 * J);
    std::memcpy(&F[FP][static_cast<size_t>(slot) * J], R.p, 8 * J);
    std::memcpy(&F[FU][static_cast<size_t>(slot) * J], R.u, 8 * J);
    std::memcpy(&F[FV][static_cast<size_t>(slot) * J], R.v, 8 * J);
    std::memcpy(&F[FW][static_cast<size_t>(slot) * J], R.w, 8 * J);
    sx[slot] = *R.x; sdx[slot] = *R.dx; sxc[slot] = *R.xc; sxvol[slot] = *R.xvol;
    sz[slot] = *R.z; sdz[slot] = *R.dz; szc[slot] = *R.zc; szvol[slot] = *R.zvol;
    srot_v[slot] = *R.rot_v;
    svbg[slot] = R.v[0];   // ghost rings are uniform; interior rings get val[].v below
    if (interior) {
      sres_v[slot] = *R.res_v;
      srot_j[slot] = *R.rot_j;
      sinc_j[slot] = *R.inc_j;
      if (*R.rot_j < 0 || *R.rot_j >= J) return fail(c, PN_ERR_LAYOUT, "ring %ld: rot_j %d outside [0, %d)", idx, *R.rot_j, J);
    }
    // every ring must carry the phi geometry of ring 0 (init.cpp:201-206, 316-319)
    RingView R0 = ring_view(const_cast<void*>(rings), J, 0);
    if (std::memcmp(R.y, R0.y, 8 * J) || std::memcmp(R.dy, R0.dy, 8 * J) || std::memcmp(R.yc, R0.yc, 8 * J) ||
        std::memcmp(R.yvol, R0.yvol, 8 * J))
      return fail(c, PN_ERR_LAYOUT, "ring %ld: phi geometry differs from ring 0", idx);
    return PN_OK;
  };
  int rc;
  for (int k = 0; k < K; ++k)
    for (int i = 0; i < I; ++i) {
      const long m = i + static_cast<long>(I) * k;
      RingView R = ring_view(const_cast<void*>(rings), J, m);
      if (*R.i != i || *R.k != k) return fail(c, PN_ERR_LAYOUT, "ring %ld carries (i,k) = (%d,%d), expected (%d,%d)", m, *R.i, *R.k, i, k);
      const int s = slot_of(d, i, k);
      if ((rc = take(rings, m, s, true)) != PN_OK) return rc;
      svbg[s] = val[m].v;
    }
  // ghost rings keep rot_j = 0, inc_j = 0, res_v = 0 (never rotated, never shifted)
  for (int k = 0; k < K; ++k)
    for (int i = 0; i < PN_NPAD; ++i) {
      if ((rc = take(lft, i + PN_NPAD * k, slot_of(d, i - PN_NPAD, k), false)) != PN_OK) return rc;
      if ((rc = take(rgh, i + PN_NPAD * k, slot_of(d, I + i, k), false)) != PN_OK) return rc;
    }
  if (c->P.ndim == 3)
    for (int i = 0; i < I; ++i)
      for (int kk = 0; kk < PN_NPAD; ++kk) {
        if ((rc = take(udr, kk + PN_NPAD * i, slot_of(d, i, kk - PN_NPAD), false)) != PN_OK) return rc;
        if ((rc = take(top, kk + PN_NPAD * i, slot_of(d, i, K + kk), false)) != PN_OK) return rc;
      }
  // is w carried?  It is identically zero in a 2-D state built by the reference's init; a
  // non-zero w is advected passively like any other field
  int has_w = (c->P.ndim == 3) ? 1 : 0;
  if (!has_w)
    for (double wv : F[FW])
      if (wv != 0.0 || std::signbit(wv)) {
        has_w = 1;
        break;
      }
  if (is_f32(c) && has_w && c->P.ndim == 2)
    return fail(c, PN_ERR_UNSUPPORTED, "the 2-D FP32 path carries no w (2-D disk with w = 0)");
  c->dp.has_w = has_w;
  // hoisted per-ring closures, evaluated on the host in FP64 exactly as the reference's
  // device code evaluates them per cell (ppm.cu:81-82, 139; disk_profile.cu:39-42, 149-152)
  Disk D{c->P};
  std::vector<double> sradc(nslot, 1.0), scs2(nslot, 0.0), snu(nslot, 0.0);
  for (long s = 0; s < nslot; ++s) {
    double rad = (c->P.ngeomx > 0) ? sxc[s] : 1.0;
    double pol = (c->P.ndim == 3) ? szc[s] : 0.0;
    double rc_ = rad;
    if (c->P.ngeomz == 5) rc_ *= std::sin(pol);
    sradc[s] = rc_;
    scs2[s] = D.cs2(rc_);
    snu[s] = D.nu(rc_);
  }
  RingView R0 = ring_view(const_cast<void*>(rings), J, 0);
  c->h_y.assign(R0.y, R0.y + J);
  c->h_dy.assign(R0.dy, R0.dy + J);
  c->h_yc.assign(R0.yc, R0.yc + J);
  c->h_yvol.assign(R0.yvol, R0.yvol + J);
  c->h_x = sx; c->h_dx = sdx; c->h_xc = sxc; c->h_xvol = sxvol;
  c->h_z = sz; c->h_dz = sdz; c->h_zc = szc; c->h_zvol = szvol;
  c->h_vbg = svbg;
  c->h_cs2 = scs2;

  hipStream_t st = c->stream;
  const size_t cells = static_cast<size_t>(nslot) * J;
  if (!is_f32(c)) {
    for (int b = 0; b < 2; ++b)
      for (int f = 0; f < kNField; ++f)
        PN_HIP(c, hipMemcpyAsync(d.f[b][f], F[f].data(), cells * sizeof(double), hipMemcpyHostToDevice, st));
  } else {
    const bool d3 = (c->P.ndim == 3);
    const bool adia = (c->P.eos == 3);
    const int nf = (d3 || adia) ? 4 : 3;
    // backgrounds: the SymDisk record of an interior ring, a ghost ring's own (uniform) values
    std::vector<double> rbg(nslot);
    for (long s = 0; s < nslot; ++s) rbg[s] = F[FR][static_cast<size_t>(s) * J];
    for (int k = 0; k < K; ++k)
      for (int i = 0; i < I; ++i) rbg[slot_of(d, i, k)] = val[i + static_cast<long>(I) * k].r;
    // EOS 3: the background pressure of the same records, and q_b = p_b / Sigma_b
    std::vector<double> pbg(nslot), qbg(nslot);
    for (long s = 0; s < nslot; ++s) pbg[s] = adia ? F[FP][static_cast<size_t>(s) * J] : scs2[s] * rbg[s];
    if (adia)
      for (int k = 0; k < K; ++k)
        for (int i = 0; i < I; ++i) pbg[slot_of(d, i, k)] = val[i + static_cast<long>(I) * k].p;
    // a ghost ring whose values a radial boundary code copies from an interior ring each step
    // (set_boundx, boundary.cu:86-206) carries that ring's backgrounds: Sigma_b and P_b (so
    // c_s^2 = P / Sigma as the copies make it), and v_bg unless the code leaves v_phi fixed
    // (code 4).  A copied deviation is then the copied value.
    std::vector<double> bcs2 = scs2, bvbg = svbg;
    for (int side = 0; side < 2; ++side) {
      const int code = side == 0 ? c->P.nlft : c->P.nrgh;
      for (int k = 0; k < K; ++k)
        for (int t = 1; t <= PN_NPAD; ++t) {
          const int si = ghost_source(code, side == 0, t, I);
          if (si < 0) continue;
          const int sg = slot_of(d, side == 0 ? -t : I - 1 + t, k), ss = slot_of(d, si, k);
          rbg[sg] = rbg[ss];
          pbg[sg] = pbg[ss];
          bcs2[sg] = bcs2[ss];
          if (code != 4) bvbg[sg] = bvbg[ss];
        }
    }
    c->h_rbg = rbg;
    for (long s = 0; s < nslot; ++s) qbg[s] = pbg[s] / rbg[s];
    c->h_qbg = qbg;
    // Sigma - Sigma_b, v_r, v_phi - v_bg and (3-D) v_theta, rounded once
    std::vector<float> G[4];
    for (int f = 0; f < nf; ++f) G[f].resize(cells);
    for (long s = 0; s < nslot; ++s)
      for (int j = 0; j < J; ++j) {
        const size_t a = static_cast<size_t>(s) * J + j;
        G[0][a] = static_cast<float>(F[FR][a] - rbg[s]);
        G[1][a] = static_cast<float>(F[FU][a]);
        G[2][a] = static_cast<float>(F[FV][a] - bvbg[s]);
        if (d3) G[3][a] = static_cast<float>(F[FW][a]);
        if (adia) G[3][a] = static_cast<float>(F[FP][a] / F[FR][a] - qbg[s]);
      }
    for (int b = 0; b < 2; ++b)
      for (int f = 0; f < nf; ++f)
        PN_HIP(c, hipMemcpyAsync(c->df.f[b][f], G[f].data(), cells * sizeof(float), hipMemcpyHostToDevice, st));
    // the per-slot table, entry-major, every entry evaluated in FP64 and rounded once; the
    // overhang past the last slot repeats it.  Radial neighbours are slots s +- 1 of the same
    // theta row, theta neighbours s +- sr.
    const long ts = c->df.tstride;
    const long sr = d.sr;
    std::vector<float> tab(static_cast<size_t>(kTabEntries) * ts, 0.0f);
    std::vector<double> pb(nslot), Lb(nslot);
    for (long s = 0; s < nslot; ++s) {
      pb[s] = pbg[s];
      Lb[s] = sradc[s] * bvbg[s];
    }
    for (long s = 0; s < nslot; ++s) {
      const long ir = s % sr, kz = s / sr;
      const bool hn = ir + 1 < sr, hp = ir > 0;
      const bool hnz = d3 && kz + 1 < d.sk, hpz = d3 && kz > 0;
      const double rad = sxc[s], xa0 = sx[s], w0 = sdx[s];
      const double th = d3 ? szc[s] : 0.5 * kPiD;
      const double st_ = d3 ? std::sin(th) : 1.0, ct_ = d3 ? std::cos(th) : 0.0;
      const double sh = std::sin(0.5 * (0.5 * kPiD - th));
      const double oms = 2.0 * sh * sh;   // 1 - sin(theta) without cancellation
      const double x1 = hn ? sx[s + 1] : xa0 + w0;
      // radial balance of the background: -1/r^2 + L_b^2 / (r^3 sin^2) - pressure
      const double V0 = d3 ? 0.5 * (xa0 * xa0 + x1 * x1) : 0.5 * (xa0 + x1);
      const double dpbm = hp ? pb[s] - pb[s - 1] : 0.0;
      const double abg = dpbm * V0 / (rbg[s] * sxvol[s]);
      const double fb = -1.0 / rad / rad + Lb[s] * Lb[s] / rad / rad / rad / (st_ * st_);
      const double Lf = hp ? 0.5 * (Lb[s] + Lb[s - 1]) : Lb[s];
      const double ffb = -1.0 / xa0 / xa0 + Lf * Lf / xa0 / xa0 / xa0 / (st_ * st_);
      double q[kTabEntries];
      for (double& v : q) v = 0.0;
      q[kTabRad] = rad;
      q[kTabXa0] = xa0;
      q[kTabDx0] = w0;
      q[kTabIxc] = 1.0 / rad;
      q[kTabRb] = rbg[s];
      q[kTabDrb] = hn ? rbg[s + 1] - rbg[s] : 0.0;
      q[kTabPb] = pb[s];
      q[kTabDpb] = hn ? pb[s + 1] - pb[s] : 0.0;
      q[kTabCs2] = bcs2[s];
      q[kTabCs] = std::sqrt(bcs2[s]);
      q[kTabLb] = Lb[s];
      q[kTabDLb] = hn ? Lb[s + 1] - Lb[s] : 0.0;
      q[kTabResBg] = fb - abg;
      q[kTabAbg] = abg;
      q[kTabDpbm] = dpbm;
      q[kTabFface] = ffb;
      q[kTabLf] = Lf;
      q[kTabIxa3] = 1.0 / (xa0 * xa0 * xa0) / (st_ * st_);
      q[kTabVrot] = sres_v[s] / sradc[s];
      q[kTabRotC] = 1.0;   // no drift until the first step's pre-step rewrites the rotation rows
      q[kTabRotPhS] = std::sin(-0.25 * c->h_dy[0]);
      q[kTabRotPhC] = std::cos(-0.25 * c->h_dy[0]);
      q[kTabDvol0] = sxvol[s];
      q[kTabQb] = qbg[s];
      q[kTabDqb] = hn ? qbg[s + 1] - qbg[s] : 0.0;
      q[kTabRc] = sradc[s];
      q[kTabIrc] = 1.0 / sradc[s];
      q[kTabIs2] = 1.0 / (st_ * st_);
      q[kTabIxa3w] = 1.0 / (xa0 * xa0 * xa0);
      q[kTabSth] = st_;
      q[kTabCth] = ct_;
      q[kTabOms] = oms;
      // the viscosity's quadratic fit (viscosity.cu:2-19) through the left edges of s-1, s,
      // s+1, applied to the backgrounds nu rho_b and Omega_b = L_b / r^2
      q[kTabNu] = snu[s];
      if (hp && hn) {
        auto fit = [&](double um, double u0, double up, double& first, double& second) {
          const double x2 = sx[s] - sx[s - 1], x3 = sx[s + 1] - sx[s - 1];
          const double h2 = (u0 - um) / x2, h3 = (up - um) / x3;
          const double A = (h3 - h2) / (x3 - x2), B = (x3 * h2 - x2 * h3) / (x3 - x2);
          second = 2.0 * A;
          first = second * x2 + B;
        };
        auto om = [&](long t) { return Lb[t] / (sradc[t] * sradc[t]); };
        double d1, d2;
        fit(snu[s - 1] * rbg[s - 1], snu[s] * rbg[s], snu[s + 1] * rbg[s + 1], d1, d2);
        q[kTabDnb] = d1;
        fit(om(s - 1), om(s), om(s + 1), d1, d2);
        q[kTabD1Ob] = d1;
        q[kTabD2Ob] = d2;
      }
      // paraset (parabola.cu:138-176) on the row's widths; neighbours past the row repeat
      // its end widths (those coefficients never reach a kept parabola)
      {
        auto w = [&](long t) { return sdx[kz * sr + std::min(std::max(t - kz * sr, 0L), sr - 1)]; };
        auto an = [&](long t) { return w(t) + w(t + 1); };
        auto bn = [&](long t) { return an(t) + w(t); };
        auto cc = [&](long t) { return an(t) + w(t + 1); };
        const double dd = 1.0 / (an(s - 1) + an(s + 1));
        q[kTabP0] = w(s) / an(s) * (1.0 + 2.0 * dd * w(s + 1) * (an(s - 1) / bn(s) - an(s + 1) / cc(s)));
        q[kTabP1] = -dd * w(s) * an(s - 1) / bn(s);
        q[kTabP2] = dd * w(s + 1) * an(s + 1) / cc(s);
        const double e = w(s) / (an(s - 1) + w(s + 1));
        q[kTabP3] = e * bn(s - 1) / an(s);
        q[kTabP4] = e * cc(s) / an(s - 1);
      }
      if (d3) {
        // theta: edges z(s), z(s) + dz(s); the background's balance of the centrifugal
        // theta component L_b^2 cos / (r sin)^3 and the theta pressure gradient
        const double za = sz[s], zb = sz[s] + sdz[s];
        const double sa = std::sin(za), ca = std::cos(za), sb = std::sin(zb), cb = std::cos(zb);
        const double dvz = szvol[s] * rad;
        const double dpbmz = hpz ? pb[s] - pb[s - sr] : 0.0;
        const double abgz = dpbmz * 0.5 * (sa + sb) / (rbg[s] * dvz);
        const double r3 = rad * rad * rad;
        const double Gc = ct_ / (r3 * st_ * st_ * st_);
        const double Lfz = hpz ? 0.5 * (Lb[s] + Lb[s - sr]) : Lb[s];
        const double Gf = ca / (r3 * sa * sa * sa);
        q[kTabZa0] = za;
        q[kTabDz0] = sdz[s];
        q[kTabDvolZ] = dvz;
        q[kTabSzl] = sa;
        q[kTabCzl] = ca;
        q[kTabSzr] = sb;
        q[kTabCzr] = cb;
        q[kTabDrbZ] = hnz ? rbg[s + sr] - rbg[s] : 0.0;
        q[kTabDpbZ] = hnz ? pb[s + sr] - pb[s] : 0.0;
        q[kTabDLbZ] = hnz ? Lb[s + sr] - Lb[s] : 0.0;
        q[kTabDpbmZ] = dpbmz;
        q[kTabResBgZ] = Lb[s] * Lb[s] * Gc - abgz;
        q[kTabAbgZ] = abgz;
        q[kTabFfaceZ] = Lfz * Lfz * Gf;
        q[kTabLfZ] = Lfz;
        q[kTabGfZ] = Gf;
        q[kTabGcZ] = Gc;
        const double is = 1.0 / st_;
        q[kTabGc1Z] = (-is * is - 3.0 * ct_ * ct_ * is * is * is * is) / r3;
        q[kTabGc2Z] = (8.0 * ct_ * is * is * is + 12.0 * ct_ * ct_ * ct_ * is * is * is * is * is) / r3;
        // the theta viscosity's fit through the left edges r z(s - sr), r z(s), r z(s + sr)
        if (hpz && hnz) {
          const double x2 = rad * (sz[s] - sz[s - sr]), x3 = rad * (sz[s + sr] - sz[s - sr]);
          auto fitz = [&](double um, double u0, double up, double& first, double& second) {
            const double h2 = (u0 - um) / x2, h3 = (up - um) / x3;
            const double A = (h3 - h2) / (x3 - x2), B = (x3 * h2 - x2 * h3) / (x3 - x2);
            second = 2.0 * A;
            first = second * x2 + B;
          };
          double d1, d2;
          const double nv = c->P.vis_nu;
          fitz(nv * rbg[s - sr], nv * rbg[s], nv * rbg[s + sr], d1, d2);
          q[kTabDnbZ] = d1;
          fitz(bvbg[s - sr], bvbg[s], bvbg[s + sr], d1, d2);
          q[kTabD1VZ] = d1;
          q[kTabD2VZ] = d2;
        }
      }
      for (int e = 0; e < kTabEntries; ++e) {
        tab[e * ts + s] = static_cast<float>(q[e]);
        if (s == nslot - 1)
          for (long o = nslot; o < ts; ++o) tab[e * ts + o] = static_cast<float>(q[e]);
      }
    }
    std::vector<float> t_dy(J), t_yvol(J);
    for (int j = 0; j < J; ++j) {
      t_dy[j] = static_cast<float>(c->h_dy[j]);
      t_yvol[j] = static_cast<float>(c->h_yvol[j]);
    }
    PN_HIP(c, hipMemcpyAsync(c->df.tab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice, st));
    std::vector<float> tabz;
    if (d3) {
      // theta-major copy: slot s = ir + sr kz goes to kz + sk ir; the overhang repeats the last
      tabz.assign(tab.size(), 0.0f);
      for (int e = 0; e < kTabEntries; ++e) {
        for (long s = 0; s < nslot; ++s) tabz[e * ts + (s / sr) + d.sk * (s % sr)] = tab[e * ts + s];
        for (long o = nslot; o < ts; ++o) tabz[e * ts + o] = tab[e * ts + nslot - 1];
      }
      PN_HIP(c, hipMemcpyAsync(c->df.tabz, tabz.data(), tabz.size() * sizeof(float), hipMemcpyHostToDevice, st));
    }
    PN_HIP(c, hipMemcpyAsync(const_cast<float*>(c->df.dy), t_dy.data(), J * sizeof(float), hipMemcpyHostToDevice, st));
    PN_HIP(c, hipMemcpyAsync(const_cast<float*>(c->df.yvol), t_yvol.data(), J * sizeof(float), hipMemcpyHostToDevice, st));
    PN_HIP(c, hipMemcpyAsync(const_cast<double*>(c->df.rbg), c->h_rbg.data(), nslot * sizeof(double), hipMemcpyHostToDevice, st));
    PN_HIP(c, hipMemcpyAsync(const_cast<double*>(c->df.vbg), bvbg.data(), nslot * sizeof(double), hipMemcpyHostToDevice, st));
    c->df.nwin_x = f32_windows(I, kF32VariantsX[c->var_x][0]);
    c->df.nwin_y = f32_windows(J, kF32VariantsY[c->var_y][0]);
    c->df.nwin_z = f32_windows(K, kF32VariantsZ[c->var_z][0]);
    c->df.uniform_phi = 1;
    for (int j = 1; j < J; ++j)
      if (c->h_dy[j] != c->h_dy[0]) c->df.uniform_phi = 0;
    c->df.uniform_theta = d3 ? 1 : 0;
    for (long s = 1; d3 && s < nslot; ++s)
      if (sdz[s] != sdz[0]) c->df.uniform_theta = 0;
    c->df.pack_tails = c->pack_tails;
    // the host vectors above must outlive the copies
    PN_HIP(c, hipStreamSynchronize(st));
  }
  auto up = [&](const void* dst, const void* src, size_t n) {
    return hipMemcpyAsync(const_cast<void*>(dst), src, n, hipMemcpyHostToDevice, st);
  };
  const size_t sb = nslot * sizeof(double);
  PN_HIP(c, up(d.x, sx.data(), sb));
  PN_HIP(c, up(d.dx, sdx.data(), sb));
  PN_HIP(c, up(d.xc, sxc.data(), sb));
  PN_HIP(c, up(d.xvol, sxvol.data(), sb));
  PN_HIP(c, up(d.z, sz.data(), sb));
  PN_HIP(c, up(d.dz, sdz.data(), sb));
  PN_HIP(c, up(d.zc, szc.data(), sb));
  PN_HIP(c, up(d.zvol, szvol.data(), sb));
  PN_HIP(c, up(d.radc, sradc.data(), sb));
  PN_HIP(c, up(d.cs2, scs2.data(), sb));
  PN_HIP(c, up(d.nu, snu.data(), sb));
  PN_HIP(c, up(d.rot_v, srot_v.data(), sb));
  PN_HIP(c, up(d.res_v, sres_v.data(), sb));
  PN_HIP(c, up(d.rot_j, srot_j.data(), nslot * sizeof(int32_t)));
  PN_HIP(c, up(d.inc_j, sinc_j.data(), nslot * sizeof(int32_t)));
  PN_HIP(c, up(d.y, c->h_y.data(), J * sizeof(double)));
  PN_HIP(c, up(d.dy, c->h_dy.data(), J * sizeof(double)));
  PN_HIP(c, up(d.yc, c->h_yc.data(), J * sizeof(double)));
  PN_HIP(c, up(d.yvol, c->h_yvol.data(), J * sizeof(double)));
  PN_HIP(c, up(d.val, val, nring * sizeof(pn_symdisk)));
  rc = sync(c);
  if (rc != PN_OK) return rc;
  if (lft != c->h_lft.data()) {
    c->h_lft.assign(static_cast<const char*>(lft), static_cast<const char*>(lft) + lr_bytes);
    c->h_rgh.assign(static_cast<const char*>(rgh), static_cast<const char*>(rgh) + lr_bytes);
    if (c->P.ndim == 3) {
      c->h_udr.assign(static_cast<const char*>(udr), static_cast<const char*>(udr) + ut_bytes);
      c->h_top.assign(static_cast<const char*>(top), static_cast<const char*>(top) + ut_bytes);
    }
    c->h_val.assign(val, val + nval);
  }
  c->cur = 0;
  c->loaded = true;
  if (c->gexec) {
    hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  return PN_OK;
}

int pn_store_rings(pn_ctx* c, void* rings, size_t rings_bytes) {
  if (!c) return PN_ERR_ARG;
  if (!c->loaded) return fail(c, PN_ERR_STATE, "no state loaded");
  const int I = c->g.imax, J = c->g.jmax, K = c->g.kmax;
  const size_t rb = pn_ring_bytes(J);
  if (!rings || rings_bytes != rb * I * K) return fail(c, PN_ERR_ARG, "rings: %zu bytes, expected %zu", rings_bytes, rb * I * K);
  const DevGrid& d = c->dg;
  const long nslot = static_cast<long>(d.sr) * d.sk;
  const size_t cells = static_cast<size_t>(nslot) * J;
  std::vector<double> F[kNField];
  for (int f = 0; f < kNField; ++f) F[f].assign(cells, 0.0);
  std::vector<float> G[4];
  const int nf = (c->P.ndim == 3 || c->P.eos == 3) ? 4 : 3;
  if (!is_f32(c)) {
    for (int f = 0; f < kNField; ++f)
      PN_HIP(c, hipMemcpyAsync(F[f].data(), d.f[c->cur][f], cells * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  } else {
    for (int f = 0; f < nf; ++f) {
      G[f].resize(cells);
      PN_HIP(c, hipMemcpyAsync(G[f].data(), c->df.f[c->cur][f], cells * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
  }
  std::vector<double> rot_v(nslot), res_v(nslot);
  std::vector<int32_t> rot_j(nslot), inc_j(nslot);
  PN_HIP(c, hipMemcpyAsync(rot_v.data(), d.rot_v, nslot * 8, hipMemcpyDeviceToHost, c->stream));
  PN_HIP(c, hipMemcpyAsync(res_v.data(), d.res_v, nslot * 8, hipMemcpyDeviceToHost, c->stream));
  PN_HIP(c, hipMemcpyAsync(rot_j.data(), d.rot_j, nslot * 4, hipMemcpyDeviceToHost, c->stream));
  PN_HIP(c, hipMemcpyAsync(inc_j.data(), d.inc_j, nslot * 4, hipMemcpyDeviceToHost, c->stream));
  int rc = sync(c);
  if (rc != PN_OK) return rc;
  if (is_f32(c))
    for (long s = 0; s < nslot; ++s)
      for (int j = 0; j < J; ++j) {
        const size_t a = static_cast<size_t>(s) * J + j;
        const double r = c->h_rbg[s] + static_cast<double>(G[0][a]);
        F[FR][a] = r;
        // isothermal closure (ppm.cu:139); EOS 3 carries q = p / Sigma
        F[FP][a] = (c->P.eos == 3) ? r * (c->h_qbg[s] + static_cast<double>(G[3][a])) : r * c->h_cs2[s];
        F[FU][a] = static_cast<double>(G[1][a]);
        F[FV][a] = c->h_vbg[s] + static_cast<double>(G[2][a]);
        if (c->P.ndim == 3) F[FW][a] = static_cast<double>(G[3][a]);
      }
  std::memset(rings, 0, rb * I * K);
  for (int k = 0; k < K; ++k)
    for (int i = 0; i < I; ++i) {
      const long m = i + static_cast<long>(I) * k;
      const int s = slot_of(d, i, k);
      RingView R = ring_view(rings, J, m);
      std::memcpy(R.r, &F[FR][static_cast<size_t>(s) * J], 8 * J);
      std::memcpy(R.p, &F[FP][static_cast<size_t>(s) * J], 8 * J);
      std::memcpy(R.u, &F[FU][static_cast<size_t>(s) * J], 8 * J);
      std::memcpy(R.v, &F[FV][static_cast<size_t>(s) * J], 8 * J);
      std::memcpy(R.w, &F[FW][static_cast<size_t>(s) * J], 8 * J);
      *R.x = c->h_x[s]; *R.dx = c->h_dx[s]; *R.xvol = c->h_xvol[s]; *R.xc = c->h_xc[s];
      std::memcpy(R.y, c->h_y.data(), 8 * J);
      std::memcpy(R.dy, c->h_dy.data(), 8 * J);
      std::memcpy(R.yvol, c->h_yvol.data(), 8 * J);
      std::memcpy(R.yc, c->h_yc.data(), 8 * J);
      *R.z = c->h_z[s]; *R.dz = c->h_dz[s]; *R.zvol = c->h_zvol[s]; *R.zc = c->h_zc[s];
      *R.i = i;
      *R.k = k;
      *R.rot_v = rot_v[s];
      *R.res_v = res_v[s];
      *R.rot_j = rot_j[s];
      *R.inc_j = inc_j[s];
    }
  return PN_OK;
}

int pn_get_dt(pn_ctx* c, double dt_prev, double FrRot, double* dt) {
  if (!c || !dt) return PN_ERR_ARG;
  if (!c->loaded) return fail(c, PN_ERR_STATE, "no state loaded");
  StepScalars h{};
  int rc = download_scalars(c, &h);
  if (rc != PN_OK) return rc;
  h.dt = dt_prev;
  h.FrRot = FrRot;
  h.bad = 0;
  if ((rc = upload_scalars(c, h)) != PN_OK) return rc;
  if ((rc = launch_cfl_any(c, c->cur)) != PN_OK) return rc;
  if ((rc = download_scalars(c, &h)) != PN_OK) return rc;
  *dt = h.dt;
  if (h.bad) return fail(c, PN_ERR_NUMERIC, "non-finite or non-positive dt %g", h.dt);
  return PN_OK;
}

int pn_index_shift(pn_ctx* c, double dt, double FrRot) {
  if (!c) return PN_ERR_ARG;
  if (!c->loaded) return fail(c, PN_ERR_STATE, "no state loaded");
  if (!(dt > 0.0)) return fail(c, PN_ERR_ARG, "dt %g", dt);
  StepScalars h{};
  int rc = download_scalars(c, &h);
  if (rc != PN_OK) return rc;
  h.dt = dt;
  h.FrRot = FrRot;
  if ((rc = upload_scalars(c, h)) != PN_OK) return rc;
  PN_LAUNCH(c, "index_shift", launch_index_shift(lctx(c)));
  return sync(c);
}

int pn_sweep2d(pn_ctx* c, double dt, const pn_body* planet, double FrRot) {
  if (!c || !planet) return PN_ERR_ARG;
  if (!c->loaded) return fail(c, PN_ERR_STATE, "no state loaded");
  if (!(dt > 0.0)) return fail(c, PN_ERR_ARG, "dt %g", dt);
  StepScalars h{};
  int rc = download_scalars(c, &h);
  if (rc != PN_OK) return rc;
  h.dt = dt;
  h.FrRot = FrRot;
  h.planet = *planet;
  if ((rc = upload_scalars(c, h)) != PN_OK) return rc;
  int cur = c->cur;
  if ((rc = launch_sweeps(c, cur, 0, true, true)) != PN_OK) return rc;
  if ((rc = sync(c)) != PN_OK) return rc;
  c->cur = cur;
  return PN_OK;
}

int pn_sweep_axis(pn_ctx* c, int axis, double dt, const pn_body* planet, double FrRot) {
  if (!c || !planet) return PN_ERR_ARG;
  if (!c->loaded) return fail(c, PN_ERR_STATE, "no state loaded");
  if (!(dt > 0.0)) return fail(c, PN_ERR_ARG, "dt %g", dt);
  if (axis < 0 || axis > 2 || (axis == 2 && c->P.ndim != 3)) return fail(c, PN_ERR_ARG, "axis %d", axis);
  StepScalars h{};
  int rc = download_scalars(c, &h);
  if (rc != PN_OK) return rc;
  h.dt = dt;
  h.FrRot = FrRot;
  h.planet = *planet;
  if ((rc = upload_scalars(c, h)) != PN_OK) return rc;
  LaunchCtx L = lctx(c);
  if (is_f32(c)) {
    PN_LAUNCH(c, "prep", f32_prep(L, c->df));
    if (axis == 0) PN_LAUNCH(c, "ghosts", f32_fill_ghosts(L, c->df, c->cur));
    if (axis == 0) PN_LAUNCH(c, "sweep_x", f32_x(c, L, c->var_x, c->cur));
    else if (axis == 1) PN_LAUNCH(c, "sweep_y", f32_sweep_y(L, c->df, c->var_y, c->cur, 0));
    else PN_LAUNCH(c, "sweep_z", f32_sweep_z(L, c->df, c->var_z, c->cur, 0));
  } else {
    if (axis == 0) PN_LAUNCH(c, "ghosts", launch_fill_ghosts(L, c->cur));
    hipError_t e = (axis == 0) ? fp64_sweep_x(L, c->cur) : (axis == 1) ? fp64_sweep_y(L, c->cur) : fp64_sweep_z(L, c->cur);
    if (e != hipSuccess) return fail(c, PN_ERR_HIP, "sweep %d: %s", axis, hipGetErrorString(e));
  }
  if ((rc = sync(c)) != PN_OK) return rc;
  c->cur ^= 1;
  return PN_OK;
}

int pn_time_kernels(pn_ctx* c, int nrep, double* ms6) {
  if (!c || !ms6 || nrep < 1) return PN_ERR_ARG;
  if (!c->loaded) return fail(c, PN_ERR_STATE, "no state loaded");
  for (int q = 0; q < 6; ++q) ms6[q] = 0.0;
  LaunchCtx L = lctx(c);
  int rc = PN_OK;
  for (int q = 0; q < 6 && rc == PN_OK; ++q) {
    if (q == 4 && c->P.ndim != 3) continue;
    if (q == 5 && c->P.fargo_flag == 0) continue;
    rc = event_time(c, c->stream, nrep, &ms6[q], [&]() -> hipError_t {
      hipError_t e = hipSuccess;
      switch (q) {
        case 0:
          if (is_f32(c)) {
            int n = 0;
            e = f32_cfl(L, c->df, c->cur, &n);
            if (e == hipSuccess) e = launch_dt_finalize(L, n);
          } else {
            e = launch_cfl(L, c->cur);
          }
          break;
        case 1: e = launch_index_shift(L); break;
        case 2:
          e = is_f32(c) ? f32_x(c, L, c->var_x, c->cur) : fp64_sweep_x(L, c->cur);
          c->cur ^= 1;
          break;
        case 3:
          e = is_f32(c) ? f32_sweep_y(L, c->df, c->var_y, c->cur, c->P.ndim == 2) : fp64_sweep_y(L, c->cur);
          c->cur ^= 1;
          break;
        case 4:
          e = is_f32(c) ? f32_sweep_z(L, c->df, c->var_z, c->cur, 1) : fp64_sweep_z(L, c->cur);
          c->cur ^= 1;
          break;
        case 5: e = launch_rotate(L); break;
      }
      return e;
    });
  }
  return rc;
}

int pn_damp_edges(pn_ctx* c, double dt) {
  if (!c) return PN_ERR_ARG;
  if (!c->loaded) return fail(c, PN_ERR_STATE, "no state loaded");
  if (!(dt > 0.0)) return fail(c, PN_ERR_ARG, "dt %g", dt);
  StepScalars h{};
  int rc = download_scalars(c, &h);
  if (rc != PN_OK) return rc;
  h.dt = dt;
  if ((rc = upload_scalars(c, h)) != PN_OK) return rc;
  LaunchCtx L = lctx(c);
  PN_LAUNCH(c, "damp", is_f32(c) ? f32_damp(L, c->df, c->cur) : launch_damp(L, c->cur));
  return sync(c);
}

int pn_copy_bandwidth(size_t nbytes, int reps, double* gbs) {
  if (!gbs || reps < 1 || nbytes < (1u << 20) || nbytes % 16) return PN_ERR_ARG;
  *gbs = 0.0;
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, nbytes);
  if (e == hipSuccess) e = hipMalloc(&b, nbytes);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipMemsetAsync(a, 0, nbytes, st);
  if (e == hipSuccess) e = launch_copy(st, a, b, nbytes);
  if (e == hipSuccess) e = hipEventRecord(e0, st);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy(st, a, b, nbytes);
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess && ms > 0.0f) *gbs = 2.0 * static_cast<double>(nbytes) * reps / (ms * 1e-3) / 1e9;
  if (e1) hipEventDestroy(e1);
  if (e0) hipEventDestroy(e0);
  if (st) hipStreamDestroy(st);
  if (b) hipFree(b);
  if (a) hipFree(a);
  return e == hipSuccess ? PN_OK : PN_ERR_HIP;
}

int pn_tune(pn_ctx* c, int nrep, int* var_x, int* var_y) {
  if (!c || nrep < 1) return PN_ERR_ARG;
  if (!c->loaded) return fail(c, PN_ERR_STATE, "no state loaded");
  if (!is_f32(c)) {
    if (var_x) *var_x = 0;
    if (var_y) *var_y = 0;
    return PN_OK;
  }
  LaunchCtx L = lctx(c);
  PN_LAUNCH(c, "prep", f32_prep(L, c->df));
  // every candidate reads the current buffer and writes the other one, so the state is
  // left as it was
  double best_x = 1e30, best_y = 1e30, best_z = 1e30;
  int bx = c->var_x, by = c->var_y, bz = c->var_z;
  for (int v = 0; v < kNumF32VariantsZ; ++v) {
    double tx = 0.0, ty = 0.0, tz = 0.0;
    int rc = PN_OK;
    const bool xy = v < kNumF32Variants;
    if (xy) rc = event_time(c, c->stream, nrep, &tx, [&]() { return f32_x(c, L, v, c->cur); });
    const bool z_ok = c->P.ndim == 3 && f32_z_variant_allowed(v, c->P.visc_flag != 0);
    if (rc == PN_OK && z_ok) {
      // the theta sweep's windows follow its own variant
      DevF32 fz = c->df;
      fz.nwin_z = f32_windows(c->g.kmax, kF32VariantsZ[v][0]);
      rc = event_time(c, c->stream, nrep, &tz, [&]() { return f32_sweep_z(L, fz, v, c->cur, 0); });
    }
    const bool y_ok = xy && f32_y_variant_allowed(v, c->P.visc_flag != 0);
    if (rc == PN_OK && y_ok)
      rc = event_time(c, c->stream, nrep, &ty, [&]() { return f32_sweep_y(L, c->df, v, c->cur, 0); });
    if (rc != PN_OK) return rc;
    if (xy && tx < best_x) { best_x = tx; bx = v; }
    if (y_ok && ty < best_y) { best_y = ty; by = v; }
    if (z_ok && tz < best_z) { best_z = tz; bz = v; }
  }
  c->var_x = bx;
  c->var_y = by;
  c->var_z = bz;
  c->df.nwin_x = f32_windows(c->g.imax, kF32VariantsX[bx][0]);
  c->df.nwin_y = f32_windows(c->g.jmax, kF32VariantsY[by][0]);
  c->df.nwin_z = f32_windows(c->g.kmax, kF32VariantsZ[bz][0]);
  c->tuned = true;
  if (c->gexec) {
    hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  if (var_x) *var_x = bx;
  if (var_y) *var_y = by;
  return PN_OK;
}

int pn_get_variants(const pn_ctx* c, int* var_x, int* var_y) {
  if (!c || !var_x || !var_y) return PN_ERR_ARG;
  *var_x = is_f32(c) ? c->var_x : 0;
  *var_y = is_f32(c) ? c->var_y : 0;
  return PN_OK;
}

int pn_set_variant_z(pn_ctx* c, int var_z) {
  if (!c) return PN_ERR_ARG;
  if (!is_f32(c) || c->P.ndim != 3) return var_z == 0 ? PN_OK : fail(c, PN_ERR_ARG, "no FP32 theta sweep here");
  if (var_z < 0 || var_z >= kNumF32VariantsZ)
    return fail(c, PN_ERR_ARG, "theta variant %d outside [0, %d)", var_z, kNumF32VariantsZ);
  if (!f32_z_variant_allowed(var_z, c->P.visc_flag != 0))
    return fail(c, PN_ERR_ARG, "theta variant %d: a viscous disk keeps 128-cell theta windows", var_z);
  c->var_z = var_z;
  c->df.nwin_z = f32_windows(c->g.kmax, kF32VariantsZ[var_z][0]);
  c->tuned = true;
  if (c->gexec) {
    hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  return PN_OK;
}

int pn_get_variant_z(const pn_ctx* c, int* var_z) {
  if (!c || !var_z) return PN_ERR_ARG;
  *var_z = (is_f32(c) && c->P.ndim == 3) ? c->var_z : 0;
  return PN_OK;
}

int pn_set_packing(pn_ctx* c, int on) {
  if (!c || (on != 0 && on != 1)) return PN_ERR_ARG;
  c->pack_tails = on;
  c->df.pack_tails = on;
  if (c->gexec) {   // the captured launches carry the old flag
    hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  return PN_OK;
}

int pn_set_variants(pn_ctx* c, int var_x, int var_y) {
  if (!c) return PN_ERR_ARG;
  if (!is_f32(c)) return (var_x == 0 && var_y == 0) ? PN_OK : fail(c, PN_ERR_ARG, "the FP64 path has one variant");
  if (var_x < 0 || var_x >= kNumF32Variants || var_y < 0 || var_y >= kNumF32Variants)
    return fail(c, PN_ERR_ARG, "variants (%d, %d) outside [0, %d)", var_x, var_y, kNumF32Variants);
  if (!f32_y_variant_allowed(var_y, c->P.visc_flag != 0))
    return fail(c, PN_ERR_ARG, "azimuthal variant %d: a viscous disk keeps 128-cell azimuthal windows", var_y);
  c->var_x = var_x;
  c->var_y = var_y;
  c->var_z = var_x;   // the theta sweep takes the radial shape (pn_set_variant_z sets its own)
  c->df.nwin_x = f32_windows(c->g.imax, kF32VariantsX[var_x][0]);
  c->df.nwin_y = f32_windows(c->g.jmax, kF32VariantsY[var_y][0]);
  c->df.nwin_z = f32_windows(c->g.kmax, kF32VariantsZ[var_x][0]);
  c->tuned = true;
  if (c->gexec) {
    hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  return PN_OK;
}

int pn_reduce(pn_ctx* c, const pn_body* planet, double exclude, int mode, double* out) {
  if (!c || !planet || !out) return PN_ERR_ARG;
  if (!c->loaded) return fail(c, PN_ERR_STATE, "no state loaded");
  if (mode != 0 && mode != 8 && mode != 9) return fail(c, PN_ERR_ARG, "mode %d", mode);
  StepScalars h{};
  int rc = download_scalars(c, &h);
  if (rc != PN_OK) return rc;
  h.planet = *planet;
  if ((rc = upload_scalars(c, h)) != PN_OK) return rc;
  if (is_f32(c)) PN_LAUNCH(c, "reduce", f32_reduce(lctx(c), c->df, c->cur, mode, exclude * c->P.sc_h, 3));
  else PN_LAUNCH(c, "reduce", launch_reduce(lctx(c), c->cur, mode, exclude * c->P.sc_h, 3));
  if ((rc = download_scalars(c, &h)) != PN_OK) return rc;
  *out = h.red[3];
  return PN_OK;
}

int pn_totals(pn_ctx* c, double* mass, double* angmom) {
  if (!c || !mass || !angmom) return PN_ERR_ARG;
  if (!c->loaded) return fail(c, PN_ERR_STATE, "no state loaded");
  if (is_f32(c)) PN_LAUNCH(c, "totals", f32_totals(lctx(c), c->df, c->cur, c->out2));
  else PN_LAUNCH(c, "totals", launch_totals(lctx(c), c->cur, c->out2));
  double h[2];
  PN_HIP(c, hipMemcpyAsync(h, c->out2, sizeof h, hipMemcpyDeviceToHost, c->stream));
  int rc = sync(c);
  if (rc != PN_OK) return rc;
  *mass = h[0];
  *angmom = h[1];
  return PN_OK;
}

int pn_totals_noalloc(pn_ctx* c, double* mass, double* angmom) {
  if (!c || !mass || !angmom) return PN_ERR_ARG;
  if (!c->loaded) return fail(c, PN_ERR_STATE, "no state loaded");
  double* scr = nullptr;
  size_t nscr = 0;
  PN_HIP(c, noalloc_scratch(&scr, &nscr));
  // the totals kernels run at most kMaxPartials / 2 grid-stride blocks of two partials each,
  // then the two sums: every grid fits the array
  if (nscr < static_cast<size_t>(kMaxPartials) + 2)
    return fail(c, PN_ERR_ARG, "scratch of %zu doubles holds fewer than the %d partials and 2 sums", nscr,
                kMaxPartials);
  LaunchCtx L = lctx(c);
  L.partials = scr;
  double* sums = scr + kMaxPartials;
  if (is_f32(c)) PN_LAUNCH(c, "totals", f32_totals(L, c->df, c->cur, sums));
  else PN_LAUNCH(c, "totals", launch_totals(L, c->cur, sums));
  double h[2];
  PN_HIP(c, hipMemcpyAsync(h, sums, sizeof h, hipMemcpyDeviceToHost, c->stream));
  int rc = sync(c);
  if (rc != PN_OK) return rc;
  *mass = h[0];
  *angmom = h[1];
  return PN_OK;
}

int pn_ring_means(pn_ctx* c, int field, double* out, size_t n) {
  if (!c || !out) return PN_ERR_ARG;
  if (!c->loaded) return fail(c, PN_ERR_STATE, "no state loaded");
  if (field < 0 || field > 4) return fail(c, PN_ERR_ARG, "field %d outside 0..4", field);
  const size_t nring = static_cast<size_t>(c->g.imax) * c->g.kmax;
  if (n != nring) return fail(c, PN_ERR_ARG, "out holds %zu values, the disk has %zu rings", n, nring);
  // the state's field order (r, p, u, v, w) is the hydr_ring's
  if (is_f32(c)) PN_LAUNCH(c, "ring means", f32_ring_means(lctx(c), c->df, c->cur, field, c->ringbuf));
  else PN_LAUNCH(c, "ring means", launch_ring_means(lctx(c), c->cur, field, c->ringbuf));
  PN_HIP(c, hipMemcpyAsync(out, c->ringbuf, nring * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int pn_planet_forces(pn_ctx* c, pn_body* b, double FrRot, double simtime) {
  if (!c || !b) return PN_ERR_ARG;
  const pn_params& P = c->P;
  b->m = (simtime < P.t_growth) ? std::fmax(P.M_p * std::sin(3.1415926535897932384626433832795 * simtime / (2.0 * P.t_growth)), 0.01 * 0.000003)
                                : P.M_p;
  if (P.plnt_flag == 2) {
    double f8 = 0.0, f9 = 0.0;
    int rc = pn_reduce(c, b, 0.0, 8, &f8);
    if (rc == PN_OK) rc = pn_reduce(c, b, 0.0, 9, &f9);
    if (rc != PN_OK) return rc;
    const double Sigma_0 = 1.0 * 0.00019126835;
    b->fx = Sigma_0 * f8 / b->m - 1.0 / b->x / b->x + b->x * ((b->vy + FrRot) * (b->vy + FrRot));
    b->fy = Sigma_0 * f9 / b->m;
  } else {
    b->fx = -1.0 / b->x / b->x + b->x * ((b->vy + FrRot) * (b->vy + FrRot));
    b->fy = 0.0;
  }
  return PN_OK;
}

}  // extern "C"

namespace {

// main.cu:270-330: dt = endtime, simtime, timem = 0, a fresh planet and its forces at simtime,
// and on a fresh start the priming Index_Shift (a restart keeps the FARGO state of the rings
// it read back)
int prime_run(pn_ctx* c, double simtime, bool restart, pn_run_state* st) {
  if (!c->loaded) return fail(c, PN_ERR_STATE, "no state loaded");
  std::memset(st, 0, sizeof *st);
  st->dt = c->P.endtime;
  st->simtime = simtime;
  pn_init_planet(&c->P, &st->planet, &st->FrRot);
  int rc = pn_planet_forces(c, &st->planet, st->FrRot, st->simtime);
  if (rc != PN_OK || restart) return rc;
  return pn_index_shift(c, st->dt, st->FrRot);
}

}  // namespace

extern "C" {

int pn_run_begin(pn_ctx* c, pn_run_state* st) {
  if (!c || !st) return PN_ERR_ARG;
  return prime_run(c, 0.0, false, st);
}

int pn_run_restart(pn_ctx* c, int startat, pn_run_state* st) {
  if (!c || !st) return PN_ERR_ARG;
  if (startat < 0) return fail(c, PN_ERR_ARG, "startat %d", startat);
  return prime_run(c, c->P.tmovie * (double)startat, true, st);
}

}  // extern "C"

namespace pn {
namespace rt {
namespace {

// n steps from the current buffer: captured-graph replays for whole units of per_graph steps
// (a 3-D step flips the buffer, so its graph holds two), the rest launched directly.  A
// graph is captured from the buffer the state is in and captured again when an odd number of
// direct 3-D steps has moved the state to the other one.
int run_steps(pn_ctx* c, int nsteps, int graph) {
  const int per_graph = (c->P.ndim == 3) ? 2 : 1;
  int done = 0, rc;
  if (graph && nsteps >= per_graph) {
    if (c->gexec && c->gexec_cur != c->cur) {
      hipGraphExecDestroy(c->gexec);
      c->gexec = nullptr;
    }
    if (!c->gexec) {
      hipGraph_t gr;
      PN_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int cur = c->cur;
      rc = PN_OK;
      for (int s = 0; s < per_graph && rc == PN_OK; ++s) rc = launch_step(c, cur);
      hipError_t e = hipStreamEndCapture(c->stream, &gr);
      if (rc != PN_OK) return rc;
      if (e != hipSuccess) return fail(c, PN_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
      if (cur != c->cur) return fail(c, PN_ERR_STATE, "captured steps do not return to their buffer");
      e = hipGraphInstantiate(&c->gexec, gr, nullptr, nullptr, 0);
      hipGraphDestroy(gr);
      if (e != hipSuccess) return fail(c, PN_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
      c->gexec_cur = c->cur;
    }
    while (nsteps - done >= per_graph) {
      PN_HIP(c, hipGraphLaunch(c->gexec, c->stream));
      done += per_graph;
    }
  }
  int cur = c->cur;
  for (; done < nsteps; ++done)
    if ((rc = launch_step(c, cur)) != PN_OK) return rc;
  c->cur = cur;
  return PN_OK;
}

}  // namespace

// steps that cannot reach a landing from dt0: each dt is at most 1.1 times the one before,
// so n steps span at most dt0 (1.1 + ... + 1.1^n); room is the time to the nearer of the
// next picture and endtime (a small margin covers the rounding of the device's sums)
int steps_before_landing(double dt0, double room, int cap) {
  int n = 0;
  double t = 0.0, d = dt0;
  while (n < cap) {
    d *= 1.1;
    t += d;
    if (!(t < 0.999 * room)) break;
    ++n;
  }
  return n;
}

}  // namespace rt
}  // namespace pn

extern "C" {

int pn_run(pn_ctx* c, int nsteps, int graph, pn_run_state* st) {
  if (!c || !st || nsteps < 0) return PN_ERR_ARG;
  if (!c->loaded) return fail(c, PN_ERR_STATE, "no state loaded");
  // a landing needs time left before both marks, or its step would be cut to dt <= 0
  if (st->land && !(st->timem < c->P.tmovie && st->simtime < c->P.endtime))
    return fail(c, PN_ERR_ARG, "landing with no time left: timem %g of tmovie %g, simtime %g of endtime %g",
                st->timem, c->P.tmovie, st->simtime, c->P.endtime);
  int rc;
  if (is_f32(c) && !c->tuned && (rc = pn_tune(c, 3, nullptr, nullptr)) != PN_OK) return rc;
  StepScalars h{};
  h.dt = st->dt;
  h.FrRot = st->FrRot;
  h.simtime = st->simtime;
  h.planet = st->planet;
  h.ncycle = st->ncycle;
  h.timem = st->timem;
  h.land = st->land ? 1 : 0;
  h.event = 0;
  if ((rc = upload_scalars(c, h)) != PN_OK) return rc;
  const bool fused = fused_cfl(c);
  // with the CFL taken in the sweep epilogue, dt of the first step comes from the state
  if (fused && nsteps > 0 && (rc = launch_cfl_any(c, c->cur)) != PN_OK) return rc;
  if (!h.land) {
    if ((rc = rt::run_steps(c, nsteps, graph)) != PN_OK) return rc;
  } else {
    // the landing is applied on the device every step; the host only stops after a step
    // that landed, reading the scalars between batches that cannot reach one
    int done = 0;
    while (done < nsteps) {
      const double room = std::fmin(c->P.tmovie - h.timem, c->P.endtime - h.simtime);
      const int n = std::max(1, rt::steps_before_landing(h.dt, room, nsteps - done));
      if ((rc = rt::run_steps(c, n, graph)) != PN_OK) return rc;
      done += n;
      if ((rc = download_scalars(c, &h)) != PN_OK) return rc;
      if (h.event != 0 || h.bad) break;
    }
  }
  if ((rc = download_scalars(c, &h)) != PN_OK) return rc;
  // fused: the record already holds the next step's dt; report the last one used
  st->dt = (fused && nsteps > 0) ? h.dt_prev : h.dt;
  st->FrRot = h.FrRot;
  st->simtime = h.simtime;
  st->planet = h.planet;
  st->ncycle = h.ncycle;
  st->timem = h.timem;
  st->event = h.event;
  if (h.bad) return fail(c, PN_ERR_NUMERIC, "non-finite or non-positive dt during the run");
  return PN_OK;
}

int pn_save_snapshot(pn_ctx* c, const char* path) {
  if (!c || !path) return PN_ERR_ARG;
  const size_t n = pn_ring_bytes(c->g.jmax) * c->g.imax * c->g.kmax;
  std::vector<char> buf(n);
  int rc = pn_store_rings(c, buf.data(), n);
  if (rc != PN_OK) return rc;
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(c, PN_ERR_IO, "cannot open %s for writing", path);
  size_t w = std::fwrite(buf.data(), 1, n, f);
  int cl = std::fclose(f);
  if (w != n || cl != 0) return fail(c, PN_ERR_IO, "short write to %s", path);
  return PN_OK;
}

int pn_load_snapshot(pn_ctx* c, const char* path) {
  if (!c || !path) return PN_ERR_ARG;
  if (!c->loaded) return fail(c, PN_ERR_STATE, "ghost rings and background must be loaded first");
  const size_t n = pn_ring_bytes(c->g.jmax) * c->g.imax * c->g.kmax;
  std::vector<char> buf(n);
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(c, PN_ERR_IO, "snapshot %s does not exist", path);
  size_t r = std::fread(buf.data(), 1, n, f);
  int extra = std::fgetc(f);
  std::fclose(f);
  if (r != n) return fail(c, PN_ERR_IO, "snapshot %s holds %zu bytes, expected %zu", path, r, n);
  if (extra != EOF) return fail(c, PN_ERR_IO, "snapshot %s is longer than %zu bytes", path, n);
  const size_t lr = c->h_lft.size(), ut = c->h_udr.size();
  return pn_load_rings(c, buf.data(), n, c->h_lft.data(), c->h_rgh.data(), lr, ut ? c->h_udr.data() : nullptr,
                       ut ? c->h_top.data() : nullptr, ut, c->h_val.data(), c->h_val.size());
}

}  // extern "C"
