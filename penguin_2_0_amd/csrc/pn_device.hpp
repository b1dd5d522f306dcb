// pn_device.hpp -- device-side records shared by the HIP kernels and the host runtime.
//
// Layout in HBM (one context, one precision):
//   fields     per-field SoA, slot-major, phi-contiguous:  f[buf][field][slot*jmax + j]
//              two buffers (ping-pong): a sweep reads one and writes the other, so every
//              window sees pre-update halos, as the reference's tile carry gives
//              (ppm.cu:88-118, 219-264)
//   slot       ring index over a padded ring grid: slot = (i+6) + (imax+12)*(k+k0),
//              k0 = 6 in 3-D (theta ghost rings), 0 in 2-D; ghost rings live in the same
//              arrays, so a radial pencil is a plain walk over slots
//   per slot   x, dx, xc, xvol, z, dz, zc, zvol, rad_cyl, cs2(rad_cyl), nu(rad_cyl),
//              rot_v, res_v, rot_j, inc_j   (the scalar members of hydr_ring,
//              variable_types.h:75-96, plus hoisted per-ring closures)
//   per column y, dy, yc, yvol: one shared copy of the phi geometry (the reference keeps
//              one copy per ring, variable_types.h:80-83)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "penguin.h"

namespace pn {

constexpr int kNpad = 6;       // halo cells per side (variable_types.h:13)
constexpr int kWin = 128;      // window length of the reference sweeps (arrsize, variable_types.h:15)
constexpr int kReal = kWin - 2 * kNpad;  // interior cells per window (realarr)
constexpr int kWaveSize = 64;

enum Field { FR = 0, FP = 1, FU = 2, FV = 3, FW = 4, kNField = 5 };

struct DevGrid {
  int imax, jmax, kmax, ndim;
  int sr;        // slots along r: imax + 12
  int sk;        // slots along theta: kmax + 12 (3-D) or 1
  int k0;        // theta offset of interior rings in slot space
  int nwin_x, nwin_y, nwin_z;   // windows per pencil
  double* f[2][kNField];
  const double *x, *dx, *xc, *xvol, *z, *dz, *zc, *zvol;
  const double *radc, *cs2, *nu;
  double *rot_v, *res_v;
  int *rot_j, *inc_j;
  const double *y, *dy, *yc, *yvol;
  const pn_symdisk* val;
};

struct DevParams {
  double gam, courant, sc_h, p_alpha, p_beta, vis_nu;
  double M_p, t_growth, rs_fac;
  int eos, flat, plnt, visc, fargo, frrot, ngeomx, ngeomy, ngeomz, ndim, has_w;
  int nlft, nrgh, kill, bary;
  double xmin, xmax;
  double endtime, tmovie;
};

// Scalars one step reads: device-resident so a whole step can be replayed as a graph.
struct StepScalars {
  double dt;          // dt of the current step
  double dt_prev;     // dt of the previous step (input of the CFL cap)
  double FrRot;
  double simtime;
  pn_body planet;
  double cfl_max;     // max signal speed / width behind the last dt
  double red[4];      // reduction results (mass, modes 8 / 9)
  int64_t ncycle;
  int bad;            // non-finite dt seen
  int trig_ok;        // FP32 planet trig table holds the planet azimuth trig_y
  double trig_y;
  double fr_shift;    // FrRot before this step's opening half kick (what Index_Shift reads)
  double timem;       // time since the last picture (main.cu:273)
  double old_dt;      // the CFL dt a landing step cut (main.cu:354, 362)
  int land;           // apply the endtime / tmovie landing (main.cu:352-366)
  int event;          // the landing this step made: 0, 1 tmovie, 2 endtime
};

__host__ __device__ inline int slot_of(const DevGrid& g, int i, int k) {
  return (i + kNpad) + g.sr * (k + g.k0);
}

// set_boundx (boundary.cu:86-206): the interior ring whose values radial ghost ring t
// (1..kNpad cells outside the edge, on the left or right) takes under boundary code
// 0 (mirrored, v_r negated), 1 and 4 (the edge ring) or 2 (periodic); code 3 keeps the
// ghost ring fixed (-1)
__host__ __device__ inline int ghost_source(int code, bool left, int t, int imax) {
  switch (code) {
    case 0: return left ? t : imax - 1 - t;
    case 1:
    case 4: return left ? 0 : imax - 1;
    case 2: return left ? imax - t : t - 1;
    default: return -1;
  }
}

}  // namespace pn
