// pn_f32.hpp -- device records and launchers of the bulk-FP32 path.
#pragma once

#include "pn_launch.hpp"

namespace pn {

// per-slot constants of the FP32 sweeps, entry-major: entry e of slot s at tab[e*stride + s]
// (FP64 values rounded once; rho_b is Sigma_b, p_b = c_s^2 Sigma_b (EOS 3: the initial
// pressure of the background record), L_b = r v_bg):
//   rad (cell centre), xa0 (left edge), dx0 (width), 1/rad,
//   rho_b, rho_b(s+1) - rho_b(s), p_b, p_b(s+1) - p_b(s),
//   c_s^2, c_s, L_b, L_b(s+1) - L_b(s),
//   res_bg, a_bg, dp_bm = p_b(s) - p_b(s-1), f_face,
//   L_f = (L_b(s) + L_b(s-1))/2, 1/xa0^3, vrot = res_v / r (rewritten each step), dvol0,
//   nu, (nu rho_b)', Omega_b', Omega_b'', and c0..c4 of paraset (parabola.cu:138-176) on the
//   slot widths
// with, in FP64: a_bg = dp_bm V0 / (rho_b dvol0), V0 = (xa0(s) + xa0(s+1))/2, the pressure
// acceleration of the background; res_bg = -1/rad^2 + L_b^2/rad^3 - a_bg, the background's
// net radial acceleration; f_face = -1/xa0^2 + L_f^2/xa0^3, the force at the left edge;
// Omega_b = L_b / rad^2, and the derivatives are those of the viscosity's quadratic fit
// through the left edges of slots s-1, s, s+1 (viscosity.cu:2-19).
// In 3-D (spherical r, theta) rad is r, L_b = r_cyl v_bg is the angular momentum about the
// axis, the centrifugal factor of L_b at the left edge is 1/(xa0^3 sin^2 theta) (kTabIxa3),
// the star's gravity of the background is -1/r^2, and the theta sweep has its own set:
//   theta left edge, width, volume r (cos - cos), sin / cos of the left and right edges,
//   background differences to the next theta slot (s + stride in slots), the theta
//   residual and pressure accelerations, the edge force, L_f, and G = cos / (r sin)^3 of
//   the centrifugal theta component at the edge and at the centre with two derivatives.
// The stride leaves room past the last slot for a radial window's overhang (copies of the
// last slot, as a clamped read would give).
enum TabEntry {
  kTabRad, kTabXa0, kTabDx0, kTabIxc,
  kTabRb, kTabDrb, kTabPb, kTabDpb,
  kTabCs2, kTabCs, kTabLb, kTabDLb,
  kTabResBg, kTabAbg, kTabDpbm, kTabFface,
  kTabLf, kTabIxa3, kTabVrot, kTabDvol0,
  kTabNu, kTabDnb, kTabD1Ob, kTabD2Ob,
  kTabP0, kTabP1, kTabP2, kTabP3, kTabP4,   // paraset coefficients of the Eulerian radial grid
  // r_cyl, 1/r_cyl, 1/sin^2(theta), 1/xa0^3, sin / cos theta, 1 - sin theta (all 1, 1/r... in 2-D)
  kTabRc, kTabIrc, kTabIs2, kTabIxa3w, kTabSth, kTabCth, kTabOms,
  // theta sweep
  kTabZa0, kTabDz0, kTabDvolZ, kTabSzl, kTabCzl, kTabSzr, kTabCzr,
  kTabDrbZ, kTabDpbZ, kTabDLbZ, kTabDpbmZ,
  kTabResBgZ, kTabAbgZ, kTabFfaceZ, kTabLfZ, kTabGfZ, kTabGcZ, kTabGc1Z, kTabGc2Z,
  // theta viscosity (viscosity.cu:474-518): the quadratic fit's first derivative of
  // vis_nu rho_b and first and second derivatives of v_bg along x = r theta
  kTabDnbZ, kTabD1VZ, kTabD2VZ,
  // per step: sin / cos of dt (vrot - v_y) / 4, half the angle the planet's azimuth moves
  // against the slot's frame in half a step (rewritten with vrot)
  kTabRotS, kTabRotC,
  // per step, azimuthal sweep on a uniform grid: the same from the column's left edge,
  // sin / cos of (dt (vrot - v_y) / 2 - dphi / 2) / 2
  kTabRotPhS, kTabRotPhC,
  // EOS 3 (2-D): q_b = p_b / rho_b and q_b(s+1) - q_b(s)
  kTabQb, kTabDqb,
  kTabEntries
};
constexpr int kTabOverhang = 256;   // >= the longest radial window

// FP32 state as deviations from per-slot backgrounds, and the per-slot constants.
// Fields: 0 Sigma - Sigma_b(slot), 1 v_r, 2 v_phi - v_bg(slot), 3 v_theta (3-D) or
// p / Sigma - q_b(slot) (2-D, EOS 3).
struct DevF32 {
  float* f[2][4];
  float* tab;                  // kTabEntries x tstride, entry-major
  int tstride;                 // nslot + kTabOverhang, a multiple of 4
  // 3-D: the same table theta-major (slot (ir, kz) at kz + sk ir, so a theta pencil's slots
  // are consecutive); stride tstride; its per-step rows are rewritten beside the others
  float* tabz;
  const double *rbg, *vbg;     // Sigma_b and v_bg per slot, FP64
  const float *dy, *yvol;      // per column
  float2* trig;                // per column, per step: sin / cos of half the angle to the planet
  int nwin_x, nwin_y, nwin_z;  // windows of the variant in use
  int uniform_phi;             // every dphi equal: the azimuthal paraset is (1/2, -1/6, 1/6, 1/2, 1/2)
  int uniform_theta;           // every dtheta equal (3-D)
  int pack_tails;              // radial / theta: a last window of at most 32 K cells runs two
                               // columns per wavefront (same results, half the passes)
};

constexpr int kNumF32Variants = 4;
// radial sweep: (cells per lane K, waves per workgroup W, columns per wave Q); a workgroup
// stages a window of W*Q columns through LDS
extern const int kF32VariantsX[kNumF32Variants][3];
// theta sweep: kF32VariantsX, plus (3, 4, 2): 192-cell windows, which hold two theta pencils of
// 64 cells with their halos per wavefront (76 cells each)
constexpr int kNumF32VariantsZ = 5;
extern const int kF32VariantsZ[kNumF32VariantsZ][3];
// azimuthal sweep: (cells per lane K, waves per workgroup W)
extern const int kF32VariantsY[kNumF32Variants][2];

// The reference's azimuthal viscosity (viscosity.cu:423-472) runs after the remap on the
// window's cells and reads the neighbours c +- 1, so at the edges of the kept range it sees
// halo cells the remap left Lagrangian: window boundaries are part of its discretization.
// Viscous disks therefore keep the reference's 128-cell windows (K = 2) in the azimuthal
// sweep; inviscid sweeps give the same result for any window length.
inline bool f32_y_variant_allowed(int variant, bool viscous) {
  return !viscous || kF32VariantsY[variant][0] == 2;
}
// The theta viscosity (viscosity.cu:474-518, in the theta sweep) also runs after the remap
// and reads c +- 1, so a viscous 3-D disk keeps 128-cell theta windows (K = 2) as well.
inline bool f32_z_variant_allowed(int variant, bool viscous) {
  return !viscous || kF32VariantsZ[variant][0] == 2;
}

int f32_windows(int len, int K);
// per-step tables: planet trig per column, vrot per slot (after Index_Shift)
hipError_t f32_prep(const LaunchCtx& c, const DevF32& f);
hipError_t f32_sweep_x(const LaunchCtx& c, const DevF32& f, int variant, int src);
// do_cfl: write each workgroup's CFL maximum of the stored state to c.cflp[workgroup]
// (the azimuthal sweep in 2-D, the theta sweep in 3-D: the step's last sweep)
hipError_t f32_sweep_y(const LaunchCtx& c, const DevF32& f, int variant, int src, int do_cfl);
int f32_cfl_partials_y(const DevGrid& g, int variant);   // partials the fused CFL writes
// 3-D: the spherical radial sweep and the theta sweep, both in the radial sweep's variants
hipError_t f32_sweep_x3(const LaunchCtx& c, const DevF32& f, int variant, int src);
hipError_t f32_sweep_z(const LaunchCtx& c, const DevF32& f, int variant, int src, int do_cfl);
int f32_cfl_partials_z(const DevGrid& g, int variant);

// batches of disks: one persistent launch per sweep over every disk's work items.
// items: workgroup-sized work items of one disk (radial: window x column group; azimuthal:
// the y partial count).  Radial: order holds nslots (disk, item) pairs (nslots a multiple of
// 8; disk -1 for an empty slot), slot t taken by the workgroup that runs it.  Azimuthal:
// base holds ndisk + 1 item offsets.  Both in device memory, as dd / dfs (per-disk records).
// grid: f32_batch_grid(axis, variant, bary, adia).  bary: the disks run in the barycentric
// frame (the kernels with its star terms; the others lack them); adia: the disks carry EOS 3
// (the kernels with the fourth field).
int f32_batch_grid(int axis, int variant, bool bary, bool adia);
hipError_t f32_sweep_x_batch(hipStream_t s, const DiskDev* dd, const DevF32* dfs, const int* order, int nslots,
                             int variant, int src, int grid, bool bary, bool adia);
hipError_t f32_sweep_y_batch(hipStream_t s, const DiskDev* dd, const DevF32* dfs, const int* base, int ndisk,
                             int variant, int src, int grid, bool bary, bool adia);
// CFL maxima per block into c.cflp; returns the partial count through *n
hipError_t f32_cfl(const LaunchCtx& c, const DevF32& f, int src, int* n);
// CFL partials f32_cfl writes for this grid
int f32_cfl_blocks(const DevGrid& g);
hipError_t f32_reduce(const LaunchCtx& c, const DevF32& f, int src, int mode, double exclude, int slot);
hipError_t f32_totals(const LaunchCtx& c, const DevF32& f, int src, double* out2);
// azimuthal mean of field (0 Sigma, 1 P, 2 v_r, 3 v_phi, 4 v_theta) per interior ring into out
hipError_t f32_ring_means(const LaunchCtx& c, const DevF32& f, int src, int field, double* out);
// damping zones at the radial edges with ss->dt (the FP64 path's launch_damp on the FP32 state)
hipError_t f32_damp(const LaunchCtx& c, const DevF32& f, int src);
// radial boundary codes other than 3 (the FP64 path's launch_fill_ghosts on the FP32 state);
// nothing is launched when both codes are 3
hipError_t f32_fill_ghosts(const LaunchCtx& c, const DevF32& f, int src);

}  // namespace pn
