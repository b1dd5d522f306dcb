// pn_kernels_fp64.hip -- FP64 directional sweeps (the drop-in path).
//
// Built with -ffp-contract=off so each operation rounds as the reference's does; the
// remaining differences to a CPU FP64 build are the last-ulp results of the device
// transcendentals (sqrt, exp, log, pow, sin, cos).
//
// Work decomposition: one wavefront per (pencil, window).  Windows are the reference's
// (ppm.cu:34-48, 169-196): 128 cells starting every 116, with 6 halo cells per side, the
// last one shortened to the pencil's remainder (a remainder of zero gets a full window
// instead of being skipped).  Four waves per 256-thread workgroup; workgroup ids are
// remapped so that consecutive logical workgroups land on the same XCD (ids are dealt
// round-robin to the 8 XCDs) and neighbouring columns share that XCD's L2.
#include "pn_launch.hpp"
#include "ppm_wave.hpp"

namespace pn {

namespace {

constexpr int kK = 2;              // cells per lane: 64 lanes x 2 = one 128-cell window
constexpr int kWavesPerBlock = 4;
constexpr double kTwoPi = 6.283185307179586476925286766559;

__device__ __forceinline__ int xcd_swizzle(int bid, int nblk) {
  const int nx = 8;
  int q = nblk / nx, rem = nblk % nx;
  int x = bid % nx, idx = bid / nx;
  return x * q + (x < rem ? x : rem) + idx;
}

__device__ __forceinline__ int wrapj(int j, int jmax) {
  if (j < 0) j += jmax;
  if (j >= jmax) j -= jmax;
  return j;
}

__device__ __forceinline__ int win_len(int len, int nwin, int l) {
  return (l < nwin - 1) ? kWin : (len - kReal * (nwin - 1) + 2 * kNpad);
}

__device__ __forceinline__ Planet load_planet(const StepScalars* ss, int bary) {
  Planet p;
  p.bary = bary;
  p.m = ss->planet.m;
  p.x = ss->planet.x;
  p.y = ss->planet.y;
  p.vx = ss->planet.vx;
  p.vy = ss->planet.vy;
  p.rs = ss->planet.rs;
  return p;
}

// Waves per SIMD the radial sweep asks for: the inviscid isothermal cylindrical sweep needs
// 172 VGPRs, two waves; at 168 it holds three with no spill and runs 1024 x 3072 in 0.248
// instead of 0.274 ms.  Every other form spills at 168 (the viscous one four accesses, the
// other closures and the 3-D sweeps tens), so they keep what the allocator gives.
template <class CF>
constexpr int x_waves() {
  return (CF::EOS == 0 && CF::GX == 1 && CF::VISC == 0) ? 3 : 1;
}

// ---------------------------------------------------------------------------------------
// radial sweep (sweepx, ppm.cu:26-160): pencil = lab column j0 of theta layer k
template <class CF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(x_waves<CF>(), 8))) void k_sweep_x(DevGrid g, DevParams prm, const StepScalars* __restrict__ ss,
                                                 int src) {
  const int lane = threadIdx.x & 63;
  const int gw = xcd_swizzle(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  const int jmax = g.jmax;
  const int j0 = gw % jmax;
  const int rest = gw / jmax;
  const int l = rest % g.nwin_x;
  const int k = rest / g.nwin_x;
  if (k >= g.kmax) return;   // wave-uniform; the kernel has no barriers

  const double dt = ss->dt, FrRot = ss->FrRot;
  const Planet pl = load_planet(ss, prm.bary);
  RunFlags rf{prm.flat, prm.plnt > 0, 1, 1, prm.ndim};
  const int lim = win_len(g.imax, g.nwin_x, l);
  rf.remap_v = 1;
  rf.remap_w = prm.has_w;
  const double* __restrict__ fr = g.f[src][FR];
  const double* __restrict__ fp = g.f[src][FP];
  const double* __restrict__ fu = g.f[src][FU];
  const double* __restrict__ fv = g.f[src][FV];
  const double* __restrict__ fw = g.f[src][FW];

  Win<double, kK> W;
  C<double, kK> radc, nu;
  long addr[kK];
  int slot[kK];
  const double azi0 = g.yc[j0];
#pragma unroll
  for (int h = 0; h < kK; ++h) {
    const int c = kK * lane + h;
    int i_ring = kReal * l - kNpad + c;
    if (c >= lim) i_ring = kReal * l - kNpad + lim - 1;   // keep loads in bounds; results unused
    const int s = slot_of(g, i_ring, k);
    slot[h] = s;
    const int j = wrapj(j0 - g.rot_j[s], jmax);
    const long a = (long)s * jmax + j;
    addr[h] = a;
    const double rad = (CF::GX > 0) ? g.xc[s] : 1.0;
    const double rc = g.radc[s];
    W.rad[h] = rad;
    W.azi[h] = azi0;
    W.pol[h] = (prm.ndim == 3) ? g.zc[s] : 0.0;
    radc[h] = rc;
    nu[h] = g.nu[s];
    W.xa0[h] = g.x[s];
    W.dx0[h] = g.dx[s];
    W.dvol0[h] = g.xvol[s];
    W.vrot[h] = g.res_v[s] / rc;
    double r = fr[a], p = fp[a], u = fu[a], v = fv[a];
    double w = prm.has_w ? fw[a] : 0.0;
    if (CF::GY > 2) v *= rc;
    if (CF::GZ == 5) w *= rad;
    W.r[h] = r; W.p[h] = p; W.u[h] = u; W.v[h] = v; W.w[h] = w;
    W.e[h] = energy_of<CF>(p, r, u, v, prm.gam);
  }

  ppmlr_wave<double, kK, CF>(W, lim, dt, FrRot, pl, rf, prm.gam);
  if constexpr (CF::VISC) visc_r<double, kK>(W, lim, dt, radc, nu, prm.ndim == 3);

  double* __restrict__ dr = g.f[src ^ 1][FR];
  double* __restrict__ dp = g.f[src ^ 1][FP];
  double* __restrict__ du = g.f[src ^ 1][FU];
  double* __restrict__ dv = g.f[src ^ 1][FV];
  double* __restrict__ dw = g.f[src ^ 1][FW];
#pragma unroll
  for (int h = 0; h < kK; ++h) {
    const int c = kK * lane + h;
    if (c >= kNpad && c < lim - kNpad) {
      const long a = addr[h];
      const double r = W.r[h];
      double p;
      if constexpr (CF::EOS == 0) p = r * g.cs2[slot[h]];
      else if constexpr (CF::EOS == 1) p = g.cs2[slot[h]] * pow(r, prm.gam) / prm.gam;
      else p = W.p[h];
      dr[a] = r;
      dp[a] = p;
      du[a] = W.u[h];
      dv[a] = (CF::GY > 2) ? W.v[h] / radc[h] : W.v[h];
      if (prm.has_w) dw[a] = (CF::GZ == 5) ? W.w[h] / W.rad[h] : W.w[h];
    }
  }
}

// ---------------------------------------------------------------------------------------
// azimuthal sweep (sweepy, ppm.cu:162-307): pencil = one periodic ring
template <class CF>
__global__ __launch_bounds__(256) void k_sweep_y(DevGrid g, DevParams prm, const StepScalars* __restrict__ ss,
                                                 int src) {
  const int lane = threadIdx.x & 63;
  const int gw = xcd_swizzle(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  const int l = gw % g.nwin_y;
  const int n = gw / g.nwin_y;
  if (n >= g.imax * g.kmax) return;
  const int i = n % g.imax, k = n / g.imax;
  const int s = slot_of(g, i, k);
  const int jmax = g.jmax;

  const double dt = ss->dt, FrRot = ss->FrRot;
  const Planet pl = load_planet(ss, prm.bary);
  // sweep slots: u = v_phi, v = physical w, w = v_r (always carried)
  RunFlags rf{prm.flat, prm.plnt > 0, prm.has_w, 1, prm.ndim};
  const int lim = win_len(jmax, g.nwin_y, l);
  const int rot = g.rot_j[s];
  const double rc = g.radc[s];
  const double pol = (prm.ndim == 3) ? g.zc[s] : 0.0;
  const double vrot = g.res_v[s] / rc;
  const long base = (long)s * jmax;
  const double* __restrict__ fr = g.f[src][FR] + base;
  const double* __restrict__ fp = g.f[src][FP] + base;
  const double* __restrict__ fu = g.f[src][FU] + base;
  const double* __restrict__ fv = g.f[src][FV] + base;
  const double* __restrict__ fw = g.f[src][FW] + base;

  Win<double, kK> W;
  int jd[kK];
#pragma unroll
  for (int h = 0; h < kK; ++h) {
    const int c = kK * lane + h;
    int jref = kReal * l - kNpad + (c < lim ? c : lim - 1);
    const int j0 = wrapj(jref, jmax);
    const int j = wrapj(j0 - rot, jmax);
    jd[h] = j;
    double azi = g.yc[j0], xa0 = g.y[j0];
    double dvol0 = g.yvol[j0];
    if (jref < 0) { xa0 -= kTwoPi; azi -= kTwoPi; }
    if (jref >= jmax) { xa0 += kTwoPi; azi += kTwoPi; }
    if (CF::GY > 2) dvol0 *= rc;
    W.azi[h] = azi;
    W.xa0[h] = xa0;
    W.dx0[h] = g.dy[j0];
    W.dvol0[h] = dvol0;
    W.rad[h] = rc;
    W.pol[h] = pol;
    W.vrot[h] = vrot;
    const double r = fr[j], p = fp[j], u = fv[j];
    const double v = prm.has_w ? fw[j] : 0.0;
    W.r[h] = r; W.p[h] = p; W.u[h] = u; W.v[h] = v; W.w[h] = fu[j];
    W.e[h] = energy_of<CF>(p, r, u, v, prm.gam);
  }

  ppmlr_wave<double, kK, CF>(W, lim, dt, FrRot, pl, rf, prm.gam);
  if constexpr (CF::VISC) {
    C<double, kK> x;
#pragma unroll
    for (int h = 0; h < kK; ++h) x[h] = W.xa0[h] * rc;
    visc_p<double, kK>(W, lim, dt, rc, g.nu[s], x, prm.ndim == 3);
  }

  double* __restrict__ dr = g.f[src ^ 1][FR] + base;
  double* __restrict__ dp = g.f[src ^ 1][FP] + base;
  double* __restrict__ du = g.f[src ^ 1][FU] + base;
  double* __restrict__ dv = g.f[src ^ 1][FV] + base;
  double* __restrict__ dw = g.f[src ^ 1][FW] + base;
  const double cs2 = g.cs2[s];
#pragma unroll
  for (int h = 0; h < kK; ++h) {
    const int c = kK * lane + h;
    if (c >= kNpad && c < lim - kNpad) {
      const int j = jd[h];
      const double r = W.r[h];
      double p;
      if constexpr (CF::EOS == 0) p = r * cs2;
      else if constexpr (CF::EOS == 1) p = cs2 * pow(r, prm.gam) / prm.gam;
      else p = W.p[h];
      dr[j] = r;
      dp[j] = p;
      du[j] = W.w[h];
      dv[j] = W.u[h];
      if (prm.has_w) dw[j] = W.v[h];
    }
  }
}

// ---------------------------------------------------------------------------------------
// theta sweep (3-D): pencil = (ring i, lab column j0) over theta layers.  The reference has
// no caller for it (main.cu:376); it is assembled from the theta pieces it holds
// (forces.cu:298-372, states.cu:29-39, evolve.cu:26-29, 63-67, 110-124, remap.cu:50-54,
// init.cpp:415-541).  Sweep components: u <- v_theta (w), v <- v_r (u),
// w <- v_phi * r sin(theta) (the angular momentum get_fz's centrifugal term expects).  With
// viscosity the window then takes device_viscosity_z (viscosity.cu:474-518) on
// (v_r, v_phi, v_theta).
template <class CF>
__global__ __launch_bounds__(256) void k_sweep_z(DevGrid g, DevParams prm, const StepScalars* __restrict__ ss,
                                                 int src) {
  const int lane = threadIdx.x & 63;
  const int gw = xcd_swizzle(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  const int jmax = g.jmax;
  const int j0 = gw % jmax;
  const int rest = gw / jmax;
  const int l = rest % g.nwin_z;
  const int i = rest / g.nwin_z;
  if (i >= g.imax) return;

  const double dt = ss->dt, FrRot = ss->FrRot;
  const Planet pl = load_planet(ss, prm.bary);
  RunFlags rf{prm.flat, prm.plnt > 0, 1, 1, prm.ndim};
  const int lim = win_len(g.kmax, g.nwin_z, l);

  Win<double, kK> W;
  C<double, kK> radc;
  long addr[kK];
  const double azi0 = g.yc[j0];
#pragma unroll
  for (int h = 0; h < kK; ++h) {
    const int c = kK * lane + h;
    int kk = kReal * l - kNpad + (c < lim ? c : lim - 1);
    const int s = slot_of(g, i, kk);
    const int j = wrapj(j0 - g.rot_j[s], jmax);
    const long a = (long)s * jmax + j;
    addr[h] = a;
    const double rad = (CF::GX > 0) ? g.xc[s] : 1.0;
    const double rc = g.radc[s];
    W.rad[h] = rad;
    W.azi[h] = azi0;
    W.pol[h] = g.zc[s];
    radc[h] = rc;
    W.xa0[h] = g.z[s];
    W.dx0[h] = g.dz[s];
    W.dvol0[h] = (CF::GZ == 5) ? g.zvol[s] * rad : g.zvol[s];
    W.vrot[h] = g.res_v[s] / rc;
    const double r = g.f[src][FR][a], p = g.f[src][FP][a];
    const double u = g.f[src][FW][a], v = g.f[src][FU][a];
    W.r[h] = r; W.p[h] = p; W.u[h] = u; W.v[h] = v;
    W.w[h] = g.f[src][FV][a] * rc;
    W.e[h] = energy_of<CF>(p, r, u, v, prm.gam);
  }

  ppmlr_wave<double, kK, CF>(W, lim, dt, FrRot, pl, rf, prm.gam);
  // v_phi from the angular momentum the sweep carried
  C<double, kK> vphi;
#pragma unroll
  for (int h = 0; h < kK; ++h) vphi[h] = W.w[h] / radc[h];
  if constexpr (CF::VISC) {
    C<double, kK> x;
#pragma unroll
    for (int h = 0; h < kK; ++h) x[h] = W.xa0[h] * W.rad[h];
    visc_z<double, kK>(W.r, lim, dt, prm.vis_nu, x, W.v, vphi, W.u);
  }

#pragma unroll
  for (int h = 0; h < kK; ++h) {
    const int c = kK * lane + h;
    if (c >= kNpad && c < lim - kNpad) {
      const long a = addr[h];
      const int kk = kReal * l - kNpad + c;
      const int s = slot_of(g, i, kk);
      const double r = W.r[h];
      double p;
      if constexpr (CF::EOS == 0) p = r * g.cs2[s];
      else if constexpr (CF::EOS == 1) p = g.cs2[s] * pow(r, prm.gam) / prm.gam;
      else p = W.p[h];
      g.f[src ^ 1][FR][a] = r;
      g.f[src ^ 1][FP][a] = p;
      g.f[src ^ 1][FW][a] = W.u[h];
      g.f[src ^ 1][FU][a] = W.v[h];
      g.f[src ^ 1][FV][a] = vphi[h];
    }
  }
}

template <int AX, class CF>
void launch_one(dim3 grid, dim3 block, const LaunchCtx& c, int src) {
  if constexpr (AX == 0) k_sweep_x<CF><<<grid, block, 0, c.stream>>>(c.g, c.p, c.ss, src);
  else if constexpr (AX == 1) k_sweep_y<CF><<<grid, block, 0, c.stream>>>(c.g, c.p, c.ss, src);
  else k_sweep_z<CF><<<grid, block, 0, c.stream>>>(c.g, c.p, c.ss, src);
}

template <int AX>
hipError_t dispatch(const LaunchCtx& c, int src, long waves) {
  const int blocks = (int)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
  const DevParams& p = c.p;
  dim3 grid(blocks), block(64 * kWavesPerBlock);
#define PN_LAUNCH(GX, GY, GZ, EOS, VISC) launch_one<AX, Cfg<AX, GX, GY, GZ, EOS, VISC>>(grid, block, c, src)
  if (p.ngeomx == 1 && p.ngeomy == 3 && p.ngeomz == 0) {
    switch (p.eos * 2 + (p.visc ? 1 : 0)) {
      case 0: PN_LAUNCH(1, 3, 0, 0, 0); break;
      case 1: PN_LAUNCH(1, 3, 0, 0, 1); break;
      case 2: PN_LAUNCH(1, 3, 0, 1, 0); break;
      case 3: PN_LAUNCH(1, 3, 0, 1, 1); break;
      case 4: PN_LAUNCH(1, 3, 0, 2, 0); break;
      case 5: PN_LAUNCH(1, 3, 0, 2, 1); break;
      case 6: PN_LAUNCH(1, 3, 0, 3, 0); break;
      case 7: PN_LAUNCH(1, 3, 0, 3, 1); break;
      default: return hipErrorInvalidValue;
    }
  } else if (p.ngeomx == 2 && p.ngeomy == 4 && p.ngeomz == 5) {
    switch (p.eos * 2 + (p.visc ? 1 : 0)) {
      case 0: PN_LAUNCH(2, 4, 5, 0, 0); break;
      case 1: PN_LAUNCH(2, 4, 5, 0, 1); break;
      case 4: PN_LAUNCH(2, 4, 5, 2, 0); break;
      case 5: PN_LAUNCH(2, 4, 5, 2, 1); break;
      default: return hipErrorInvalidValue;
    }
  } else {
    return hipErrorInvalidValue;
  }
#undef PN_LAUNCH
  return hipGetLastError();
}

}  // namespace

hipError_t fp64_sweep_x(const LaunchCtx& c, int src) {
  return dispatch<0>(c, src, (long)c.g.jmax * c.g.nwin_x * c.g.kmax);
}
hipError_t fp64_sweep_y(const LaunchCtx& c, int src) {
  return dispatch<1>(c, src, (long)c.g.nwin_y * c.g.imax * c.g.kmax);
}
hipError_t fp64_sweep_z(const LaunchCtx& c, int src) {
  return dispatch<2>(c, src, (long)c.g.jmax * c.g.nwin_z * c.g.imax);
}

}  // namespace pn
