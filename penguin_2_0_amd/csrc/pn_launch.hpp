// pn_launch.hpp -- host-callable launchers for the HIP kernels (one set per precision).
#pragma once

#include <hip/hip_runtime.h>

#include "pn_device.hpp"

namespace pn {

struct LaunchCtx {
  DevGrid g;
  DevParams p;
  StepScalars* ss;      // device scalars
  double* partials;     // reduction scratch, >= kMaxPartials doubles
  double* cflp;         // CFL maxima per block or per wave, ncflp doubles
  int ncflp;
  hipStream_t stream;
};

constexpr int kMaxPartials = 8192;

// one disk of a batch, as the batched kernels read it (an array of these in device memory)
struct DiskDev {
  DevGrid g;
  DevParams p;
  StepScalars* ss;
  double* cflp;      // this disk's CFL partials
  int ncfl_post;     // partials the post-step reduces into the next dt (fused CFL), or 0
  int ncfl_pre;      // partials the pre-step reduces into this step's dt (a planet that moves:
                     // the CFL pass runs before the step, as FrRot changes inside it), or 0
  float2* trig;      // FP32 per-step tables, or null
  float* tab;        // the FP32 per-slot table (per-step rows rewritten by the pre-step), or null
  int tstride;
};

// FP64 path (pn_kernels_fp64.hip); each returns hipSuccess or the launch error
hipError_t fp64_sweep_x(const LaunchCtx& c, int src);
hipError_t fp64_sweep_y(const LaunchCtx& c, int src);
hipError_t fp64_sweep_z(const LaunchCtx& c, int src);

// common small kernels (pn_kernels_common.hip)
hipError_t launch_index_shift(const LaunchCtx& c);
hipError_t launch_rotate(const LaunchCtx& c);
hipError_t launch_cfl(const LaunchCtx& c, int src);              // partials into cflp, then ss->dt
hipError_t launch_reduce(const LaunchCtx& c, int src, int mode, double exclude, int slot);  // -> ss->red[slot]
hipError_t launch_totals(const LaunchCtx& c, int src, double* out2);  // out2: device, mass and angmom
// azimuthal mean of FP64 field `field` of every interior ring (out: device, imax kmax doubles)
hipError_t launch_ring_means(const LaunchCtx& c, int src, int field, double* out);
// simtime, drift, the disk's forces on the planet (modes 8 / 9; from the FP32 state when f32 is
// given), closing half kick
struct DevF32;
hipError_t launch_planet_after_sweep(const LaunchCtx& c, int src, const DevF32* f32 = nullptr);
hipError_t launch_dt_finalize(const LaunchCtx& c, int n);           // dt = min(1.1 dt, courant / max of n cflp)
hipError_t launch_sum_partials(const LaunchCtx& c, int nb, int stride, double* out);
// radial ghost rings of buffer src from the interior under the boundary codes (FP64 state;
// nothing for fixed ghost rings)
hipError_t launch_fill_ghosts(const LaunchCtx& c, int src);
// damping zones at the radial edges with the step's dt (FP64 state)
hipError_t launch_damp(const LaunchCtx& c, int src);
// the current device's copy of the file-scope scratch array of the no-allocation entry
// points (its address looked up once per device) and its size in doubles
hipError_t noalloc_scratch(double** ptr, size_t* ndoubles);
// plain device copy of nbytes (a multiple of 16)
hipError_t launch_copy(hipStream_t s, const void* a, void* b, size_t nbytes);
// CFL maxima per block into cflp (no dt); *n receives the partial count
hipError_t launch_cfl_partials(const LaunchCtx& c, int src, int* n);
// one workgroup before the sweeps: dt from ncfl partials (0: keep ss->dt), Index_Shift,
// opening half kick; trig / tab / tabz (FP32 tables: rows of stride tstride, tabz the 3-D
// theta-major copy) may be null.  Disks of many slots get Index_Shift and the table rows from
// a second, grid-wide kernel.
hipError_t launch_pre_step(const LaunchCtx& c, int ncfl, float2* trig, float* tab, float* tabz, int tstride);
// one workgroup after the sweeps (plnt_flag < 2): roll, time and planet update, then dt
// from ncfl partials when ncfl > 0 (disks of many slots: the roll from a grid-wide kernel;
// many partials: folded first by a grid-wide kernel into the reduction scratch)
hipError_t launch_post_step(const LaunchCtx& c, int ncfl);
// the same for every disk of a batch, one workgroup per disk (pre-step with ncfl 0; post-step
// with each disk's ncfl_post)
hipError_t launch_pre_step_batch(hipStream_t s, const DiskDev* dd, int ndisk);
hipError_t launch_post_step_batch(hipStream_t s, const DiskDev* dd, int ndisk);

}  // namespace pn
