// pn_kernels_common.hip -- the small per-step kernels: FARGO split and roll, CFL max,
// reductions, device-side planet update.  FP64 throughout (-ffp-contract=off).
#include <mutex>

#include "pn_f32.hpp"
#include "pn_launch.hpp"
#include "ppm_wave.hpp"

namespace pn {

// ---------------------------------------------------------------------------------------
// Scratch of the entry points that allocate nothing (pn_totals_noalloc): the per-block
// partials of a reduction and their sums in one file-scope array.  Every device the library
// runs on holds its own copy until the process ends.  On the host the array's name is only a
// stand-in: kernels and copies take the address hipGetSymbolAddress returns for the current
// device, looked up once per device (noalloc_scratch, first called from pn_create).
constexpr int kNoAllocDoubles = kMaxPartials + 8;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::once_flag g_scratch_once[kMaxDevices];
double* g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];
}  // namespace

hipError_t noalloc_scratch(double** ptr, size_t* ndoubles) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::call_once(g_scratch_once[dev], [dev]() {
    void* q = nullptr;
    g_scratch_err[dev] = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    g_scratch_ptr[dev] = (g_scratch_err[dev] == hipSuccess) ? static_cast<double*>(q) : nullptr;
  });
  if (g_scratch_err[dev] != hipSuccess) return g_scratch_err[dev];
  *ptr = g_scratch_ptr[dev];
  *ndoubles = kNoAllocDoubles;
  return hipSuccess;
}

namespace {

constexpr double kPi = 3.1415926535897932384626433832795;
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kEarthMass = 0.000003;
constexpr double kSigma0 = 1.0 * 0.00019126835;   // global.h:37 (MMSN_1AU)

// rotate_rings (ppm.cu:309-322)
__global__ void k_rotate(DevGrid g) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.imax * g.kmax) return;
  const int s = slot_of(g, n % g.imax, n / g.imax);
  int j = g.rot_j[s] + g.inc_j[s];
  if (j < 0) j += g.jmax;
  if (j >= g.jmax) j -= g.jmax;
  g.rot_j[s] = j;
}

// set_boundx (boundary.cu:86-206) on the FP64 state of buffer src: each radial ghost cell
// takes the values its boundary code names from the interior (a ghost ring's data index is
// its lab column less its own rot_j; the interior ring's, less that ring's).  Code 3 (fixed
// ghost rings) writes nothing.  Thread: one lab column of one of the 12 ghost rings of a
// theta row.
__global__ void k_fill_ghosts(DevGrid g, DevParams prm, int src) {
  const int jl = blockIdx.x * blockDim.x + threadIdx.x;
  if (jl >= g.jmax) return;
  const int gi = blockIdx.y % (2 * kNpad), k = blockIdx.y / (2 * kNpad);
  const bool left = gi < kNpad;
  const int code = left ? prm.nlft : prm.nrgh;
  const int t = left ? kNpad - gi : gi - kNpad + 1;   // cells outside the edge cell
  const int srci = ghost_source(code, left, t, g.imax);
  if (srci < 0) return;
  const int sg = left ? slot_of(g, -t, k) : slot_of(g, g.imax - 1 + t, k);
  const int ss = slot_of(g, srci, k);
  int jg = jl - g.rot_j[sg], js = jl - g.rot_j[ss];
  jg += (jg < 0) ? g.jmax : 0;
  jg -= (jg >= g.jmax) ? g.jmax : 0;
  js += (js < 0) ? g.jmax : 0;
  js -= (js >= g.jmax) ? g.jmax : 0;
  const long ag = (long)sg * g.jmax + jg, as = (long)ss * g.jmax + js;
  const int last = (code == 4) ? (left ? FV : FU) : kNField;   // fields [FR, last) are copied
  for (int f = FR; f < last; ++f) g.f[src][f][ag] = g.f[src][f][as];
  if (code == 0) g.f[src][FU][ag] = -g.f[src][FU][as];
}

// kill_wave (kill_wave.cu:3-43) on the FP64 state of buffer src with the step's dt: cells
// within sc_h/2 of xmin or xmax relax towards their ring's background; w relaxes to 0.
// Thread: one cell of an interior ring.
__global__ void k_damp(DevGrid g, DevParams prm, const StepScalars* __restrict__ ss, int src) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= g.jmax) return;
  const int n = blockIdx.y;   // ring i + imax k
  const int s = slot_of(g, n % g.imax, n / g.imax);
  const double d_kill = 0.5 * prm.sc_h;
  const double inner = prm.xmin + d_kill, outer = prm.xmax - d_kill;
  const double r = g.xc[s];
  double f;
  if (r < inner) f = (inner - r) / d_kill;
  else if (r > outer) f = (r - outer) / d_kill;
  else return;
  f *= ss->dt * pow(r, -1.5);
  const pn_symdisk V = g.val[n];
  const long a = (long)s * g.jmax + j;
  double* F[kNField] = {g.f[src][FR], g.f[src][FP], g.f[src][FU], g.f[src][FV], g.f[src][FW]};
  F[FR][a] += f * (V.r - F[FR][a]);
  F[FP][a] += f * (V.p - F[FP][a]);
  F[FU][a] += f * (V.u - F[FU][a]);
  F[FV][a] += f * (V.v - F[FV][a]);
  F[FW][a] += f * (-F[FW][a]);
}

__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// CFL signal speed per cell width (timestep.cu:12-73) over every interior cell, one maximum
// per block into cflp (a maximum is order independent, so the split is exact)
__global__ __launch_bounds__(256) void k_cfl(DevGrid g, DevParams prm, const StepScalars* __restrict__ ss,
                                             double* cflp, int src) {
  __shared__ double sh[4];
  const long ncell = (long)g.imax * g.kmax * g.jmax;
  const double FrRot = ss->FrRot;
  double m = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += (long)gridDim.x * blockDim.x) {
    const int j = (int)(t % g.jmax);
    const int n = (int)(t / g.jmax);
    const int s = slot_of(g, n % g.imax, n / g.imax);
    const long a = (long)s * g.jmax + j;
    const double rad = g.xc[s], rad_cyl = g.radc[s];
    double widthy = g.dy[j];
    if (prm.ngeomy > 2) widthy *= rad_cyl;
    const double cs = sqrt(prm.gam * g.f[src][FP][a] / g.f[src][FR][a]);
    const double t0 = (fabs(g.f[src][FU][a]) + cs) / g.dx[s];
    const double t1 = (fabs(g.f[src][FV][a] - g.rot_v[s] - FrRot * rad_cyl) + cs) / widthy;
    double v = fmax(t0, t1);
    if (prm.ndim == 3) {
      double widthz = g.dz[s];
      if (prm.ngeomz > 2) widthz *= rad;
      const double t2 = (fabs(g.f[src][FW][a]) + cs) / widthz;
      v = max3(t0, t1, t2);
    }
    m = fmax(m, v);
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) cflp[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// Azimuthal mean of field `field` of every interior ring: sum_j F[j] dy[j0] / sum_j dy[j0]
// over the ring's data columns j, lab column j0 = j + rot_j (device_func.cu:71) -- the dphi
// weighting of the reader's ring average (write_output_file, read/p_output.cpp:112-127, which
// divides by 2 pi, the sum of dphi over a full ring).  One workgroup of 256 per ring n =
// i + imax k: each thread sums its columns j = t, t + 256, ... in FP64, the wavefronts fold
// with lane shuffles, and the four wave sums are added in a fixed order.
__global__ __launch_bounds__(256) void k_ring_means(DevGrid g, int src, int field, double* __restrict__ out) {
  __shared__ double sh[8];
  const int n = blockIdx.x;
  const int s = slot_of(g, n % g.imax, n / g.imax);
  const double* __restrict__ F = g.f[src][field] + (long)s * g.jmax;
  const int rot = g.rot_j[s];
  double a = 0.0, w = 0.0;
  for (int j = threadIdx.x; j < g.jmax; j += 256) {
    int j0 = j + rot;
    if (j0 >= g.jmax) j0 -= g.jmax;
    const double d = g.dy[j0];
    a += F[j] * d;
    w += d;
  }
  a = wave_sum(a);
  w = wave_sum(w);
  if ((threadIdx.x & 63) == 0) {
    sh[threadIdx.x >> 6] = a;
    sh[4 + (threadIdx.x >> 6)] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) out[n] = ((sh[0] + sh[1]) + (sh[2] + sh[3])) / ((sh[4] + sh[5]) + (sh[6] + sh[7]));
}

// dt = min(1.1 dt_prev, courant / max)  (timestep.cu:135-141), max over n partials
__global__ __launch_bounds__(256) void k_dt_finalize(DevParams prm, StepScalars* ss, const double* __restrict__ cflp,
                                                     int n) {
  __shared__ double sh[4];
  double m = 0.0;
  for (int t = threadIdx.x; t < n; t += blockDim.x) m = fmax(m, cflp[t]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double mx = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
  ss->cfl_max = mx;
  const double dtc = prm.courant / mx;
  double dt = 1.1 * ss->dt;
  dt = fmin(dt, dtc);
  if (!(dt > 0.0) || isinf(dt)) ss->bad = 1;
  ss->dt_prev = ss->dt;
  ss->dt = dt;
}

// output_lv1 (device_func.cu:49-198): per-block sums in a fixed order
__global__ __launch_bounds__(256) void k_reduce(DevGrid g, DevParams prm, const StepScalars* __restrict__ ss,
                                                int src, int mode, double exclude, double* partials) {
  __shared__ double sh[4];
  const long ncell = (long)g.imax * g.kmax * g.jmax;
  const Planet pl{ss->planet.m, ss->planet.x, ss->planet.y, ss->planet.vx, ss->planet.vy, ss->planet.rs, prm.bary};
  double acc = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += (long)gridDim.x * blockDim.x) {
    const int j = (int)(t % g.jmax);
    const int n = (int)(t / g.jmax);
    const int s = slot_of(g, n % g.imax, n / g.imax);
    int j0 = j + g.rot_j[s];
    if (j0 < 0) j0 += g.jmax;
    if (j0 >= g.jmax) j0 -= g.jmax;
    const double r = g.f[src][FR][(long)s * g.jmax + j];
    double tmp = 0.0;
    if (mode == 0 || prm.plnt == 0) {
      tmp = r * g.xvol[s] * g.yvol[j0] * g.zvol[s];
    } else if (prm.ndim == 2 && (mode == 8 || mode == 9)) {
      const double rad = g.xc[s], phi = g.yc[j0];
      const double dd = (rad - pl.x) * (rad - pl.x) + rad * rad * (phi - pl.y) * (phi - pl.y);
      if (dd > exclude * exclude) {
        const double bg = pow(g.xc[s], -prm.p_alpha);
        if (mode == 8) {
          tmp = -grav_rad_cyl<double>(pl, rad, phi, 0.0, 0.0);
          tmp *= (r - bg) * g.xvol[s] * g.yvol[j0];
        } else {
          tmp = -grav_azi_cyl<double>(pl, rad, phi, 0.0, 0.0);
          tmp *= rad * (r - bg) * g.xvol[s] * g.yvol[j0];
        }
      }
    }
    acc += tmp;
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// mass and angular momentum with compensated per-thread accumulation
__global__ __launch_bounds__(256) void k_totals(DevGrid g, DevParams prm, int src, double* partials) {
  __shared__ double sh[8];
  const long ncell = (long)g.imax * g.kmax * g.jmax;
  double sm = 0.0, cm = 0.0, sl = 0.0, cl = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += (long)gridDim.x * blockDim.x) {
    const int j = (int)(t % g.jmax);
    const int n = (int)(t / g.jmax);
    const int s = slot_of(g, n % g.imax, n / g.imax);
    int j0 = j + g.rot_j[s];
    if (j0 >= g.jmax) j0 -= g.jmax;
    const long a = (long)s * g.jmax + j;
    const double dm = g.f[src][FR][a] * g.xvol[s] * g.yvol[j0] * g.zvol[s];
    const double dl = dm * g.f[src][FV][a] * g.radc[s];
    double y = dm - cm, tt = sm + y;
    cm = (tt - sm) - y;
    sm = tt;
    y = dl - cl;
    tt = sl + y;
    cl = (tt - sl) - y;
    sl = tt;
  }
  sm = wave_sum(sm - cm);
  sl = wave_sum(sl - cl);
  if ((threadIdx.x & 63) == 0) {
    sh[threadIdx.x >> 6] = sm;
    sh[4 + (threadIdx.x >> 6)] = sl;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    partials[2 * blockIdx.x + 1] = (sh[4] + sh[5]) + (sh[6] + sh[7]);
  }
}

__device__ __forceinline__ void neumaier_add(double& s, double& c, double x) {
  const double t = s + x;
  if (fabs(s) >= fabs(x)) c += (s - t) + x;
  else c += (x - t) + s;
  s = t;
}

constexpr int kSumThreads = 256;

// Neumaier sum of nb partials (stride st) into out, in a fixed order: thread t sums
// partials t, t + 256, ...; thread 0 then adds the 256 running sums and their
// compensations in thread order
__global__ __launch_bounds__(kSumThreads) void k_sum_partials(const double* partials, int nb, int st, double* out) {
  __shared__ double sh_s[kSumThreads], sh_c[kSumThreads];
  for (int q = 0; q < st; ++q) {
    double s = 0.0, c = 0.0;
    for (int b = threadIdx.x; b < nb; b += kSumThreads) neumaier_add(s, c, partials[(long)b * st + q]);
    sh_s[threadIdx.x] = s;
    sh_c[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      double S = 0.0, Cc = 0.0;
      for (int t = 0; t < kSumThreads; ++t) {
        neumaier_add(S, Cc, sh_s[t]);
        Cc += sh_c[t];
      }
      out[q] = S + Cc;
    }
    __syncthreads();
  }
}

__device__ double cs2_of(const DevParams& prm, double r) { return prm.sc_h * prm.sc_h * pow(r, -prm.p_beta); }

__device__ double set_M_p(const DevParams& prm, double t) {   // planet.cu:28-34
  if (t < prm.t_growth) return fmax(prm.M_p * sin(kPi * t / (2.0 * prm.t_growth)), 0.01 * kEarthMass);
  return prm.M_p;
}

// thread 0, before a step (main.cu:352-366): a step that would pass endtime or the next
// picture time is cut to land on it; old_dt keeps the dt it replaced (the second cut, when
// both apply, keeps the first cut's dt, as the reference does)
__device__ void land_step(const DevParams& prm, StepScalars* ss) {
  double dt = ss->dt;
  int ev = 0;
  if (ss->simtime + dt >= prm.endtime) {
    ss->old_dt = dt;
    dt = prm.endtime - ss->simtime;
    ev = 2;
  }
  if (ss->timem + dt >= prm.tmovie) {
    ss->old_dt = dt;
    dt = prm.tmovie - ss->timem;
    if (ev == 0) ev = 1;
  }
  ss->dt = dt;
  ss->event = ev;
}

// thread 0, after a step's closing kick (main.cu:379, 408-411): timem advances by the step's
// dt; after a landing it restarts and dt returns to the value the landing cut
__device__ void land_close(StepScalars* ss, double dt) {
  ss->timem += dt;
  if (ss->event != 0) {
    ss->timem = 0.0;
    ss->dt = ss->old_dt;
  }
}

// after the sweeps: time advance, drift, forces and the closing half kick
// (main.cu:378-387, planet.cu:54-93).  red[1], red[2] hold the mode 8 / 9 reductions
// when the planet moves.
__global__ void k_planet_drift(DevParams prm, StepScalars* ss) {
  if (threadIdx.x != 0) return;
  const double dt = ss->dt;
  ss->simtime += dt;
  ss->ncycle += 1;
  pn_body& b = ss->planet;
  const double old_rad = b.x;
  b.x += dt * b.vx;
  b.y += dt * b.vy;
  b.vy = ((b.vy + ss->FrRot) * old_rad * old_rad / b.x / b.x) - ss->FrRot;
  if (b.y < 0.0) b.y += kTwoPi;
  if (b.y >= kTwoPi) b.y -= kTwoPi;
  // get_h(x) = sqrt(c_s^2 x^3); the host computed rs for the current x, so it is only
  // re-evaluated when the planet actually moved radially
  if (b.x != old_rad) b.rs = prm.rs_fac * sqrt(cs2_of(prm, b.x) * b.x * b.x * b.x);
}

__global__ void k_planet_forces_kick(DevParams prm, StepScalars* ss, int do_kick) {
  if (threadIdx.x != 0) return;
  pn_body& b = ss->planet;
  b.m = set_M_p(prm, ss->simtime);
  if (prm.plnt == 2) {
    b.fx = kSigma0 * ss->red[1] / b.m - 1.0 / b.x / b.x + b.x * ((b.vy + ss->FrRot) * (b.vy + ss->FrRot));
    b.fy = kSigma0 * ss->red[2] / b.m;
  } else {
    b.fx = -1.0 / b.x / b.x + b.x * ((b.vy + ss->FrRot) * (b.vy + ss->FrRot));
    b.fy = 0.0;
  }
  if (do_kick) {
    const double dt = 0.5 * ss->dt;
    b.vx += dt * b.fx;
    b.vy += ss->FrRot + dt * b.fy / b.x / b.x;
    ss->FrRot = (prm.frrot == 1) ? b.vy : 0.0;
    b.vy -= ss->FrRot;
    land_close(ss, ss->dt);   // the end of the step
  }
}

constexpr int kStepThreads = 1024;
// a disk with more slots than this runs the per-ring work of the pre- and post-step over the
// whole grid (k_pre_rings, k_rotate) rather than in their single workgroup
constexpr int kSplitSlots = 2 * kStepThreads;
inline bool split_rings(const DevGrid& g) { return g.sr * g.sk > kSplitSlots; }

// max over n partials by the whole (single) workgroup; the result is valid in thread 0
__device__ double block_max_partials(const double* __restrict__ cflp, int n, double* sh) {
  double m = 0.0;
  for (int t = threadIdx.x; t < n; t += blockDim.x) m = fmax(m, cflp[t]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, sh[w]);
  return m;
}

// thread 0: dt = min(1.1 dt_prev, courant / max)  (timestep.cu:135-141)
__device__ void finalize_dt(const DevParams& prm, StepScalars* ss, double mx) {
  ss->cfl_max = mx;
  double dt = fmin(1.1 * ss->dt, prm.courant / mx);
  if (!(dt > 0.0) || isinf(dt)) ss->bad = 1;
  ss->dt_prev = ss->dt;
  ss->dt = dt;
}

// Index_Shift of one ring (orbital_advection.cu:12-43)
__device__ void index_shift_ring(DevGrid& g, const DevParams& prm, int n, double dt, double FrRot) {
  const int i = n % g.imax, k = n / g.imax;
  const int s = slot_of(g, i, k);
  const double xc = g.xc[s];
  const double widthy = xc * g.dy[0] / dt;
  double rot_v;
  if (prm.fargo == 1) rot_v = g.val[i + g.imax * k].v - FrRot * xc;
  else if (prm.fargo == 2) rot_v = (1.0 - FrRot) * xc;
  else rot_v = 0.0;
  const double jp = g.rot_v[s] / widthy;
  if (jp >= 0.0) g.inc_j[s] = (int)(jp + 0.5);
  else g.inc_j[s] = (int)(jp - 0.5);
  g.res_v[s] = floor(jp + 0.5) * widthy;
  g.rot_v[s] = rot_v;
}

// Index_Shift (orbital_advection.cu:12-43) as its own launch: one thread per interior ring,
// running the body the pre-step runs for every ring
__global__ void k_index_shift(DevGrid g, DevParams prm, const StepScalars* __restrict__ ss) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.imax * g.kmax) return;
  index_shift_ring(g, prm, n, ss->dt, ss->FrRot);
}

// the per-step rotation rows of the FP32 table for slot s (ghost slots keep res_v = 0):
// sin / cos of dt (res_v / r_cyl - v_y) / 4 with the planet velocity the sweeps will read
// (3-D: also into the theta-major copy tabz, where slot (ir, kz) sits at kz + sk ir)
__device__ __forceinline__ void rot_rows_slot(const DevGrid& g, double vy, float* tab, float* tabz, int tstride,
                                              double dt, int s) {
  const double h = 0.25 * dt * (g.res_v[s] / g.radc[s] - vy);
  const float sh = (float)sin(h), ch = (float)cos(h);
  tab[(long)kTabRotS * tstride + s] = sh;
  tab[(long)kTabRotC * tstride + s] = ch;
  const double hp = h - 0.25 * g.dy[0];
  tab[(long)kTabRotPhS * tstride + s] = (float)sin(hp);
  tab[(long)kTabRotPhC * tstride + s] = (float)cos(hp);
  if (tabz) {
    const long sz = s / g.sr + (long)g.sk * (s % g.sr);
    tabz[(long)kTabRotS * tstride + sz] = sh;
    tabz[(long)kTabRotC * tstride + sz] = ch;
    tabz[(long)kTabVrot * tstride + sz] = tab[(long)kTabVrot * tstride + s];
  }
}

// Index_Shift of the ring in slot s (if s holds one) with its vrot row, then the slot's
// rotation rows: everything per slot reads and writes that slot only
__device__ __forceinline__ void ring_slot(DevGrid& g, const DevParams& prm, double dt, double fr, double vy,
                                          float* tab, float* tabz, int tstride, int s) {
  const int i = s % g.sr - kNpad, k = s / g.sr - g.k0;
  if (i >= 0 && i < g.imax && k >= 0 && k < g.kmax) {
    index_shift_ring(g, prm, i + g.imax * k, dt, fr);
    if (tab) tab[(long)kTabVrot * tstride + s] = (float)(g.res_v[s] / g.radc[s]);
  }
  if (tab) rot_rows_slot(g, vy, tab, tabz, tstride, dt, s);
}

// Everything of a step before the sweeps, in one workgroup (main.cu:345-370): dt from ncfl
// CFL partials (ncfl 0: ss->dt is already this step's), Index_Shift of every ring, and the
// planet's opening half kick.  For FP32 contexts also the per-step tables: sin / cos of half
// the angle from the planet to each column centre (trig), and per slot of the table tab
// vrot = res_v / r_cyl and the planet's azimuthal drift against the slot in half a step
// (after the kick: the sweeps see that planet velocity).  With rings false the per-slot part
// (Index_Shift, vrot and the rotation rows) is left to k_pre_rings, which reads the FrRot
// Index_Shift needs from ss->fr_shift.
__device__ __forceinline__ void pre_step_body(DevGrid g, const DevParams& prm, StepScalars* ss,
                                              const double* __restrict__ cflp, int ncfl, float2* trig, float* tab,
                                              float* tabz, int tstride, bool rings) {
  __shared__ double sh[kStepThreads / 64];
  __shared__ double s_dt, s_fr, s_py;
  __shared__ int s_trig;
  if (ncfl > 0) {
    const double mx = block_max_partials(cflp, ncfl, sh);
    if (threadIdx.x == 0) finalize_dt(prm, ss, mx);
  }
  if (threadIdx.x == 0) {
    if (ss->land) land_step(prm, ss);
    s_dt = ss->dt;
    s_fr = ss->FrRot;
    ss->fr_shift = s_fr;
    s_py = ss->planet.y;   // the half kick below changes velocities only
    // the trig table depends on the planet's azimuth alone: refresh it when that moved
    s_trig = (trig != nullptr) && !(ss->trig_ok && ss->trig_y == s_py);
    if (s_trig) {
      ss->trig_ok = 1;
      ss->trig_y = s_py;
    }
  }
  __syncthreads();
  const double dt = s_dt, FrRot = s_fr;
  if (rings) {
    const int nring = g.imax * g.kmax;
    for (int n = threadIdx.x; n < nring; n += blockDim.x) {
      index_shift_ring(g, prm, n, dt, FrRot);
      if (tab) {
        const int s = slot_of(g, n % g.imax, n / g.imax);
        tab[(long)kTabVrot * tstride + s] = (float)(g.res_v[s] / g.radc[s]);
      }
    }
  }
  if (s_trig) {
    const double py = s_py;
    for (int j = threadIdx.x; j < g.jmax; j += blockDim.x) {
      const double a = 0.5 * (g.yc[j] - py);
      trig[j] = make_float2((float)sin(a), (float)cos(a));
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && prm.plnt > 0) {
    const double hdt = 0.5 * dt;
    pn_body& b = ss->planet;
    b.vx += hdt * b.fx;
    b.vy += FrRot + hdt * b.fy / b.x / b.x;
    ss->FrRot = (prm.frrot == 1) ? b.vy : 0.0;
    b.vy -= ss->FrRot;
  }
  if (tab && rings) {
    __syncthreads();
    const double vy = ss->planet.vy;
    // one workgroup per disk (the batched pre-step runs one per disk of the grid)
    for (int s = threadIdx.x; s < g.sr * g.sk; s += blockDim.x) rot_rows_slot(g, vy, tab, tabz, tstride, dt, s);
  }
}

__global__ __launch_bounds__(kStepThreads) void k_pre_step(DevGrid g, DevParams prm, StepScalars* ss,
                                                           const double* __restrict__ cflp, int ncfl, float2* trig,
                                                           float* tab, float* tabz, int tstride, int rings) {
  pre_step_body(g, prm, ss, cflp, ncfl, trig, tab, tabz, tstride, rings != 0);
}

// the per-slot part of the pre-step over the whole grid, one thread per slot, after
// k_pre_step with rings false (it has set dt, ss->fr_shift and the kicked planet velocity)
__global__ __launch_bounds__(256) void k_pre_rings(DevGrid g, DevParams prm, const StepScalars* ss, float* tab,
                                                   float* tabz, int tstride) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= g.sr * g.sk) return;
  ring_slot(g, prm, ss->dt, ss->fr_shift, ss->planet.vy, tab, tabz, tstride, s);
}

// one workgroup per disk of a batch (the fused CFL leaves dt already set: ncfl_pre 0; a
// planet that moves has its CFL partials from the pass before the step)
__global__ __launch_bounds__(kStepThreads) void k_pre_step_batch(const DiskDev* __restrict__ dd) {
  const DiskDev& D = dd[blockIdx.x];
  pre_step_body(D.g, D.p, D.ss, D.cflp, D.ncfl_pre, D.trig, D.tab, nullptr, D.tstride, true);
}

// Everything of a step after the sweeps for a planet that does not feel the disk
// (plnt_flag < 2), in one workgroup: the integer roll of every ring (ppm.cu:309-322), time
// advance, drift, forces and closing half kick (main.cu:378-387, planet.cu:54-93), then,
// with ncfl > 0, the next step's dt from the CFL partials the last sweep wrote.
// the integer roll of every ring (ppm.cu:309-322) by one workgroup
__device__ __forceinline__ void roll_rings(DevGrid g) {
  const int nring = g.imax * g.kmax;
  for (int n = threadIdx.x; n < nring; n += blockDim.x) {
    const int s = slot_of(g, n % g.imax, n / g.imax);
    int j = g.rot_j[s] + g.inc_j[s];
    if (j < 0) j += g.jmax;
    if (j >= g.jmax) j -= g.jmax;
    g.rot_j[s] = j;
  }
}

__device__ __forceinline__ void post_step_body(DevGrid g, const DevParams& prm, StepScalars* ss,
                                               const double* __restrict__ cflp, int ncfl, bool roll) {
  __shared__ double sh[kStepThreads / 64];
  if (roll && prm.fargo > 0) roll_rings(g);
  double mx = 0.0;
  if (ncfl > 0) mx = block_max_partials(cflp, ncfl, sh);
  if (threadIdx.x != 0) return;
  const double dt = ss->dt;
  ss->simtime += dt;
  ss->ncycle += 1;
  if (prm.plnt > 0) {
    pn_body& b = ss->planet;
    const double old_rad = b.x;
    b.x += dt * b.vx;
    b.y += dt * b.vy;
    b.vy = ((b.vy + ss->FrRot) * old_rad * old_rad / b.x / b.x) - ss->FrRot;
    if (b.y < 0.0) b.y += kTwoPi;
    if (b.y >= kTwoPi) b.y -= kTwoPi;
    if (b.x != old_rad) b.rs = prm.rs_fac * sqrt(cs2_of(prm, b.x) * b.x * b.x * b.x);
    b.m = set_M_p(prm, ss->simtime);
    b.fx = -1.0 / b.x / b.x + b.x * ((b.vy + ss->FrRot) * (b.vy + ss->FrRot));
    b.fy = 0.0;
    const double hdt = 0.5 * dt;
    b.vx += hdt * b.fx;
    b.vy += ss->FrRot + hdt * b.fy / b.x / b.x;
    ss->FrRot = (prm.frrot == 1) ? b.vy : 0.0;
    b.vy -= ss->FrRot;
  }
  land_close(ss, dt);
  if (ncfl > 0) finalize_dt(prm, ss, mx);
}

__global__ __launch_bounds__(kStepThreads) void k_post_step(DevGrid g, DevParams prm, StepScalars* ss,
                                                            const double* __restrict__ cflp, int ncfl, int roll) {
  post_step_body(g, prm, ss, cflp, ncfl, roll != 0);
}

// the largest of n CFL partials per workgroup, grid-strided: a disk whose fused CFL leaves
// more partials than kSplitPartials has them folded into kCflChunks maxima first, so the
// single-workgroup post-step reads few
constexpr int kSplitPartials = 4096, kCflChunks = 256;
__global__ __launch_bounds__(256) void k_max_chunks(const double* __restrict__ in, int n, double* __restrict__ out) {
  __shared__ double sh[4];
  double m = 0.0;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) m = fmax(m, in[t]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// one workgroup per disk of a batch; a planet that moves takes only the roll here, its time,
// drift, disk forces and closing kick following from launch_planet_after_sweep
__global__ __launch_bounds__(kStepThreads) void k_post_step_batch(const DiskDev* __restrict__ dd) {
  const DiskDev& D = dd[blockIdx.x];
  if (D.p.plnt == 2) {
    if (D.p.fargo > 0) roll_rings(D.g);
    return;
  }
  post_step_body(D.g, D.p, D.ss, D.cflp, D.ncfl_post, true);
}

int red_blocks(const DevGrid& g) {
  long ncell = (long)g.imax * g.kmax * g.jmax;
  long b = (ncell + 255) / 256;
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace

hipError_t launch_index_shift(const LaunchCtx& c) {
  int n = c.g.imax * c.g.kmax;
  k_index_shift<<<(n + 255) / 256, 256, 0, c.stream>>>(c.g, c.p, c.ss);
  return hipGetLastError();
}

hipError_t launch_fill_ghosts(const LaunchCtx& c, int src) {
  if (c.p.nlft == 3 && c.p.nrgh == 3) return hipSuccess;
  dim3 grid((c.g.jmax + 255) / 256, 2 * kNpad * c.g.kmax);
  k_fill_ghosts<<<grid, 256, 0, c.stream>>>(c.g, c.p, src);
  return hipGetLastError();
}

hipError_t launch_damp(const LaunchCtx& c, int src) {
  dim3 grid((c.g.jmax + 255) / 256, c.g.imax * c.g.kmax);
  k_damp<<<grid, 256, 0, c.stream>>>(c.g, c.p, c.ss, src);
  return hipGetLastError();
}

hipError_t launch_rotate(const LaunchCtx& c) {
  int n = c.g.imax * c.g.kmax;
  k_rotate<<<(n + 255) / 256, 256, 0, c.stream>>>(c.g);
  return hipGetLastError();
}

hipError_t launch_cfl(const LaunchCtx& c, int src) {
  const int nb = red_blocks(c.g);
  if (nb > c.ncflp) return hipErrorInvalidValue;
  k_cfl<<<nb, 256, 0, c.stream>>>(c.g, c.p, c.ss, c.cflp, src);
  k_dt_finalize<<<1, 256, 0, c.stream>>>(c.p, c.ss, c.cflp, nb);
  return hipGetLastError();
}

hipError_t launch_reduce(const LaunchCtx& c, int src, int mode, double exclude, int slot) {
  int nb = red_blocks(c.g);
  k_reduce<<<nb, 256, 0, c.stream>>>(c.g, c.p, c.ss, src, mode, exclude, c.partials);
  k_sum_partials<<<1, kSumThreads, 0, c.stream>>>(c.partials, nb, 1, &c.ss->red[slot]);
  return hipGetLastError();
}

hipError_t launch_ring_means(const LaunchCtx& c, int src, int field, double* out) {
  if (field < 0 || field >= kNField) return hipErrorInvalidValue;
  k_ring_means<<<c.g.imax * c.g.kmax, 256, 0, c.stream>>>(c.g, src, field, out);
  return hipGetLastError();
}

hipError_t launch_totals(const LaunchCtx& c, int src, double* out2) {
  int nb = red_blocks(c.g);
  if (nb > kMaxPartials / 2) nb = kMaxPartials / 2;
  k_totals<<<nb, 256, 0, c.stream>>>(c.g, c.p, src, c.partials);
  k_sum_partials<<<1, kSumThreads, 0, c.stream>>>(c.partials, nb, 2, out2);
  return hipGetLastError();
}

hipError_t launch_pre_step(const LaunchCtx& c, int ncfl, float2* trig, float* tab, float* tabz, int tstride) {
  if (ncfl < 0 || ncfl > c.ncflp) return hipErrorInvalidValue;
  const bool split = split_rings(c.g);
  k_pre_step<<<1, kStepThreads, 0, c.stream>>>(c.g, c.p, c.ss, c.cflp, ncfl, trig, tab, tabz, tstride, split ? 0 : 1);
  if (split) {
    const int n = c.g.sr * c.g.sk;
    k_pre_rings<<<(n + 255) / 256, 256, 0, c.stream>>>(c.g, c.p, c.ss, tab, tabz, tstride);
  }
  return hipGetLastError();
}

hipError_t launch_post_step(const LaunchCtx& c, int ncfl) {
  if (ncfl < 0 || ncfl > c.ncflp || c.p.plnt == 2) return hipErrorInvalidValue;
  const bool split = split_rings(c.g);
  if (split && c.p.fargo > 0) {
    const int n = c.g.imax * c.g.kmax;
    k_rotate<<<(n + 255) / 256, 256, 0, c.stream>>>(c.g);
  }
  const double* parts = c.cflp;
  int np = ncfl;
  if (ncfl > kSplitPartials) {   // the reduction scratch is free while a step runs
    k_max_chunks<<<kCflChunks, 256, 0, c.stream>>>(c.cflp, ncfl, c.partials);
    parts = c.partials;
    np = kCflChunks;
  }
  k_post_step<<<1, kStepThreads, 0, c.stream>>>(c.g, c.p, c.ss, parts, np, split ? 0 : 1);
  return hipGetLastError();
}

hipError_t launch_pre_step_batch(hipStream_t s, const DiskDev* dd, int ndisk) {
  if (ndisk < 1) return hipErrorInvalidValue;
  k_pre_step_batch<<<ndisk, kStepThreads, 0, s>>>(dd);
  return hipGetLastError();
}

hipError_t launch_post_step_batch(hipStream_t s, const DiskDev* dd, int ndisk) {
  if (ndisk < 1) return hipErrorInvalidValue;
  k_post_step_batch<<<ndisk, kStepThreads, 0, s>>>(dd);
  return hipGetLastError();
}

hipError_t launch_cfl_partials(const LaunchCtx& c, int src, int* n) {
  const int nb = red_blocks(c.g);
  if (nb > c.ncflp) return hipErrorInvalidValue;
  k_cfl<<<nb, 256, 0, c.stream>>>(c.g, c.p, c.ss, c.cflp, src);
  *n = nb;
  return hipGetLastError();
}

hipError_t launch_dt_finalize(const LaunchCtx& c, int n) {
  if (n < 1 || n > c.ncflp) return hipErrorInvalidValue;
  k_dt_finalize<<<1, 256, 0, c.stream>>>(c.p, c.ss, c.cflp, n);
  return hipGetLastError();
}

// a plain copy of n 16-byte words, grid-stride: the HBM reference of the measurements
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ a, float4* __restrict__ b, long n) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) b[i] = a[i];
}

hipError_t launch_copy(hipStream_t s, const void* a, void* b, size_t nbytes) {
  const long n = (long)(nbytes / 16);
  long nb = (n + 255) / 256;
  if (nb > 16384) nb = 16384;
  if (nb < 1) return hipErrorInvalidValue;
  k_copy<<<(unsigned)nb, 256, 0, s>>>(static_cast<const float4*>(a), static_cast<float4*>(b), n);
  return hipGetLastError();
}

hipError_t launch_sum_partials(const LaunchCtx& c, int nb, int stride, double* out) {
  k_sum_partials<<<1, kSumThreads, 0, c.stream>>>(c.partials, nb, stride, out);
  return hipGetLastError();
}

hipError_t launch_planet_after_sweep(const LaunchCtx& c, int src, const DevF32* f32) {
  k_planet_drift<<<1, 64, 0, c.stream>>>(c.p, c.ss);
  if (c.p.plnt == 2) {
    hipError_t e = f32 ? f32_reduce(c, *f32, src, 8, 0.0, 1) : launch_reduce(c, src, 8, 0.0, 1);
    if (e != hipSuccess) return e;
    e = f32 ? f32_reduce(c, *f32, src, 9, 0.0, 2) : launch_reduce(c, src, 9, 0.0, 2);
    if (e != hipSuccess) return e;
  }
  k_planet_forces_kick<<<1, 64, 0, c.stream>>>(c.p, c.ss, 1);
  return hipGetLastError();
}

}  // namespace pn
