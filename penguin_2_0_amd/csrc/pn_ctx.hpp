// pn_ctx.hpp -- the context record behind the C ABI's pn_ctx, and the runtime helpers the
// single-disk calls and the batch share (not part of the public interface).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "penguin.h"
#include "pn_f32.hpp"
#include "pn_launch.hpp"

struct pn_ctx {
  pn_grid g{};
  pn_params P{};
  int precision = PN_FP64;
  std::string err;
  pn::DevGrid dg{};
  pn::DevParams dp{};
  pn::DevF32 df{};                    // FP32 state and tables (FP32 contexts)
  int var_x = 0, var_y = 3;         // FP32 sweep variants (indices into kF32VariantsX / Y)
  int var_z = 0;                    // FP32 theta sweep variant (index into kF32VariantsZ)
  int pack_tails = 1;               // FP32 radial / theta: two short last windows per wavefront
  bool tuned = false;
  pn::StepScalars* ss = nullptr;      // device
  double* partials = nullptr;       // device
  double* cflp = nullptr;           // device, CFL partial maxima
  int ncflp = 0;
  double* out2 = nullptr;           // device, totals
  double* ringbuf = nullptr;        // device, one double per interior ring (ring means)
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  int cur = 0;                      // buffer holding the state
  bool loaded = false;
  // host copies of the per-ring values the device never changes (for packing)
  std::vector<double> h_x, h_dx, h_xc, h_xvol, h_z, h_dz, h_zc, h_zvol, h_vbg, h_rbg, h_cs2, h_qbg;
  std::vector<double> h_y, h_dy, h_yc, h_yvol;
  // graph of one step (2-D) or two steps (3-D, the buffers alternate)
  hipGraphExec_t gexec = nullptr;
  int gexec_cur = 0;   // the buffer the captured steps start from (and, whole, end on)
  // host copies of the fixed inputs, so a snapshot can be re-entered (restart)
  std::vector<char> h_lft, h_rgh, h_udr, h_top;
  std::vector<pn_symdisk> h_val;
};

namespace pn {
namespace rt {

int fail(pn_ctx* c, int code, const char* fmt, ...);
LaunchCtx lctx(pn_ctx* c);
bool is_f32(const pn_ctx* c);
bool fused_cfl(const pn_ctx* c);
int upload_scalars(pn_ctx* c, const StepScalars& h);
int download_scalars(pn_ctx* c, StepScalars* h);
int sync(pn_ctx* c);
int launch_cfl_any(pn_ctx* c, int src);
// steps from dt0 that cannot reach a landing room away (dt grows by at most 1.1 per step)
int steps_before_landing(double dt0, double room, int cap);

// mean time of nrep back-to-back launches, measured with HIP events on the context's stream
template <class F>
int event_time(pn_ctx* c, hipStream_t stream, int nrep, double* ms, F&& launch) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return fail(c, PN_ERR_HIP, "hipEventCreate");
  if (hipEventCreate(&e1) != hipSuccess) return fail(c, PN_ERR_HIP, "hipEventCreate");
  // nrep untimed launches first: after host-side idle the clocks take a few launches to
  // return to speed, which would otherwise be charged to the first variant or kernel timed
  hipError_t e = hipSuccess;
  for (int k = 0; k < nrep && e == hipSuccess; ++k) e = launch();
  if (e == hipSuccess) e = hipEventRecord(e0, stream);
  for (int k = 0; k < nrep && e == hipSuccess; ++k) e = launch();
  if (e == hipSuccess) e = hipEventRecord(e1, stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (e != hipSuccess) return fail(c, PN_ERR_HIP, "kernel timing: %s", hipGetErrorString(e));
  *ms = static_cast<double>(t) / nrep;
  return PN_OK;
}


}  // namespace rt
}  // namespace pn
