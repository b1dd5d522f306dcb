// pn_kernels_fp32.hip -- the bulk-FP32 sweeps (2-D isothermal disk) and their helpers.
//
// State in HBM: three FP32 fields per cell -- Sigma - Sigma_b, v_r and v_phi - v_bg --
// where Sigma_b and v_bg are the ring's FP64 background (the SymDisk record; a ghost ring's
// own uniform values).  P = c_s^2 Sigma is not stored (the isothermal closure the reference
// applies after every sweep, ppm.cu:139), and w is identically zero in 2-D.  Everything that
// must stay FP64 stays FP64 outside the sweeps: dt, the FARGO split and roll, the planet,
// the time, the backgrounds and residual rates, and the totals.
//
// Work decomposition as in the FP64 sweeps, with K cells per lane (window 64 K, halo 6) and
// the workgroup shape picked by the autotuner; the CFL maximum of the next step is taken in
// the azimuthal sweep's epilogue from the values it stores, one partial per workgroup.
#include <type_traits>

#include "pn_f32.hpp"
#include "ppm_wave_f32.hpp"

namespace pn {

// Occupancy.  A SIMD's 512 VGPRs per lane hold eight waves of at most 64 VGPRs, six of 80,
// five of 96 and four of 128; amdgpu_waves_per_eu asks the allocator to fit a count, spilling
// what does not fit to scratch.
// - LDS-staged pencils: the allocator left alone settles them at 99-122 (four waves).  A
//   workgroup of four waves puts one on each SIMD, so a fifth wave per SIMD is a fifth
//   workgroup per CU, which the LDS must hold too; for those shapes the kernels ask for five
//   (pencil_waves).  It costs a few spills per pencil (radial sweep 1024 x 3072: two) and took
//   that sweep from 75.1 to 73.3 us.  Workgroups of eight or sixteen waves gain nothing short
//   of six or eight waves per SIMD, so they ask for nothing.
// - Azimuthal windows of 128 cells: 66 VGPRs for a 2-D isothermal disk (seven waves; eight at
//   64 for four spills) and 86 for the batch (five waves; six at 80 for six spills).  Longer
//   windows and the 3-D and EOS 3 forms spill tens of values at the next step down, so they
//   keep what the allocator gives (a minimum of one wave asks for nothing).
#define PN_WAVES(n) __attribute__((amdgpu_waves_per_eu(n, 8)))

namespace {

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

template <int K>
__device__ __forceinline__ int win_len_k(int len, int nwin, int l) {
  constexpr int WIN = 64 * K, REAL = WIN - 2 * kNpad;
  return (l < nwin - 1) ? WIN : (len - REAL * (nwin - 1) + 2 * kNpad);
}

// the planet as the FP32 forces see it; the spherical forms use the reduced mass m/(1+m)
// sph: the spherical forms take the reduced mass m/(1+m); bary (read only by kernels built
// with BARY, which a batch runs for all its disks when one needs them): both bodies about
// their barycentre (the FP64 forms' BaryPos), with no indirect term
template <bool BARY>
__device__ __forceinline__ PlanetF planet_f(const StepScalars* ss, bool sph, int bary) {
  const double m = ss->planet.m, x = ss->planet.x;
  PlanetF p;
  p.rs2 = (float)(ss->planet.rs * ss->planet.rs);
  p.vy = (float)ss->planet.vy;
  p.bary = BARY && bary;
  if (p.bary) {
    const double stm = 1.0 / (1.0 + m);
    p.m = (float)(m * stm);
    p.x = (float)(x * stm);
    p.inv_x2 = 0.0f;
    p.stx = (float)(x * m * stm);
    p.stm = (float)stm;
  } else {
    p.m = sph ? (float)(m / (1.0 + m)) : (float)m;
    p.x = (float)x;
    p.inv_x2 = (float)(1.0 / (x * x));
    p.stx = 0.0f;
    p.stm = 1.0f;
  }
  return p;
}

__device__ __forceinline__ float wave_maxf(float m) {
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  return m;
}

// per step: sin / cos of half the angle from the planet to each lab column centre (FP64),
// each slot's vrot = res_v / r after Index_Shift and its half-step rotation rows
__global__ void k_prep(DevGrid g, DevF32 f, const StepScalars* __restrict__ ss, int nslot) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < g.jmax) {
    const double a = 0.5 * (g.yc[t] - ss->planet.y);
    f.trig[t] = make_float2((float)sin(a), (float)cos(a));
  }
  if (t < nslot) {
    f.tab[(long)kTabVrot * f.tstride + t] = (float)(g.res_v[t] / g.radc[t]);
    const double h = 0.25 * ss->dt * (g.res_v[t] / g.radc[t] - ss->planet.vy);
    f.tab[(long)kTabRotS * f.tstride + t] = (float)sin(h);
    f.tab[(long)kTabRotC * f.tstride + t] = (float)cos(h);
    f.tab[(long)kTabRotPhS * f.tstride + t] = (float)sin(h - 0.25 * g.dy[0]);
    f.tab[(long)kTabRotPhC * f.tstride + t] = (float)cos(h - 0.25 * g.dy[0]);
    if (f.tabz) {   // the theta-major copy: slot (ir, kz) at kz + sk ir
      const long sz = t / g.sr + (long)g.sk * (t % g.sr);
      f.tabz[(long)kTabVrot * f.tstride + sz] = f.tab[(long)kTabVrot * f.tstride + t];
      f.tabz[(long)kTabRotS * f.tstride + sz] = (float)sin(h);
      f.tabz[(long)kTabRotC * f.tstride + sz] = (float)cos(h);
    }
  }
}

// K consecutive floats from LDS (vector read; p is K-float aligned)
template <int K>
__device__ __forceinline__ C<float, K> lds_k(const float* p) {
  C<float, K> o;
  if constexpr (K == 2) {
    const float2 v = *reinterpret_cast<const float2*>(p);
    o[0] = v.x;
    o[1] = v.y;
  } else if constexpr (K == 4) {
    const float4 v = *reinterpret_cast<const float4*>(p);
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
  } else {
#pragma unroll
    for (int h = 0; h < K; ++h) o[h] = p[h];
  }
  return o;
}

// ---------------------------------------------------------------------------------------
// Radial and theta sweeps.  A pencil of these sweeps crosses rings (slot step 1 radially,
// one theta row of slots in theta), so its cells are not contiguous in HBM; the cells of
// neighbouring lab columns are.  A workgroup of W waves therefore takes one window (64 K
// cells with halos) of NC = W Q consecutive lab columns, stages the window's NF fields in LDS
// with loads that are contiguous along phi, runs Q columns per wave as pencils out of LDS,
// writes its results back in place, and stores the window's interior contiguously again.
// VISC: 0 none, 1 always, 2 as the disk's parameters say.
template <int K, int W, int Q, int NF>
struct XTile {
  static constexpr int WIN = 64 * K, NC = W * Q, LD = WIN + 2;
  float tile[NF][NC][LD];   // [field][column][cell]; LD even keeps K-float reads aligned
  int srot[WIN];            // rot_j of the window's slots
  float red[W];
};

// waves per SIMD a pencil kernel asks for: five where a fifth workgroup of four waves fits the
// CU's 160 KiB of LDS, else no request; none for K = 3 either (three cells per lane spill
// under the 96 registers of five waves and the 128 of four; unconstrained they take 165, three
// waves per SIMD, without spilling)
template <int K, int W, int Q, int NF>
constexpr int pencil_waves() {
  if (K == 3) return 1;
  return (W == 4 && 5 * sizeof(XTile<K, W, Q, NF>) <= 160 * 1024) ? 5 : 1;
}

// one work item: window l of the pencils of row `row` (radial: theta row k; theta: ring i),
// column group grp.  Workgroup-uniform; starts with a barrier so that a workgroup can run
// items back to back on the same LDS.  With CFL, returns in thread 0 the largest CFL term of
// the cells it stored (the last sweep of a 3-D step).
template <int K, int W, int Q, int G, int VISC, bool CFL, bool ADIA = false, bool BARY = false>
__device__ __forceinline__ float pencil_item(const DevGrid& g, const DevF32& f, const DevParams& prm,
                                             const StepScalars* __restrict__ ss, int src, int grp, int l, int row,
                                             XTile<K, W, Q, (G == kGeoRcyl && !ADIA ? 3 : 4)>& T) {
  constexpr int NF = (G == kGeoRcyl && !ADIA) ? 3 : 4;
  constexpr int WIN = 64 * K, REAL = WIN - 2 * kNpad, NC = W * Q;
  constexpr bool TH = (G == kGeoTheta);
  // staging: thread t moves the NIT cells c = t / NC + it (64 W / NC) of column t % NC, the
  // same cells on the way in and out
  constexpr int TPB = 64 * W, NIT = WIN * NC / TPB, CSTEP = TPB / NC;
  static_assert(TPB % NC == 0 && (WIN * NC) % TPB == 0, "staging map");
  auto& tile = T.tile;
  const int jmax = g.jmax;
  const int jb = grp * NC;
  const int lim = TH ? win_len_k<K>(g.kmax, f.nwin_z, l) : win_len_k<K>(g.imax, f.nwin_x, l);
  const int s0 = TH ? slot_of(g, row, REAL * l - kNpad) : slot_of(g, REAL * l - kNpad, row);
  const int step = TH ? g.sr : 1;   // slots between neighbouring cells of the pencil
  const int qs = threadIdx.x % NC, cs = threadIdx.x / NC;
  const int jl = jb + qs;
  __syncthreads();   // the previous item's stores have read srot and the tile
  for (int c = threadIdx.x; c < WIN; c += 64 * W) T.srot[c] = g.rot_j[s0 + step * (c < lim ? c : lim - 1)];
  __syncthreads();
  // the fields as buffer views with 32-bit byte offsets (a state buffer is below 2 GiB)
  const unsigned fbytes = 4u * (unsigned)(g.sr * g.sk) * (unsigned)jmax;
  Buf fin[NF];
#pragma unroll
  for (int fi = 0; fi < NF; ++fi) fin[fi] = Buf(f.f[src][fi], fbytes);
  // the tile's loads go out in chunks of CH cells per thread, each chunk's loads issued
  // before its first wait (at most 12 in flight: more raised the register count of the
  // pencils below and cost a wave of occupancy); columns past jmax (the last column group)
  // read column jmax - 1, which no wave computes and no store writes
  constexpr int CH = NF == 3 ? (NIT < 4 ? NIT : 4) : 2;
  static_assert(NIT % CH == 0, "staging chunks");
  const int jlc = jl < jmax ? jl : jmax - 1;
#pragma unroll
  for (int i0 = 0; i0 < NIT; i0 += CH) {
    float v[CH][NF];
#pragma unroll
    for (int it = 0; it < CH; ++it) {
      const int c = cs + CSTEP * (i0 + it);
      const unsigned o =
          4u * (unsigned)((s0 + step * (c < lim ? c : lim - 1)) * jmax + wrapj(jlc - T.srot[c], jmax));
#pragma unroll
      for (int fi = 0; fi < NF; ++fi) v[it][fi] = fin[fi].f(o);
    }
#pragma unroll
    for (int it = 0; it < CH; ++it) {
      const int c = cs + CSTEP * (i0 + it);
#pragma unroll
      for (int fi = 0; fi < NF; ++fi) tile[fi][qs][c] = v[it][fi];
    }
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform, in a scalar register
  const float dt = (float)ss->dt;
  const PlanetF pl = planet_f<BARY>(ss, G != kGeoRcyl, prm.bary);
  const float sqrt_gam = sqrtf((float)prm.gam);
  // radial: overhang cells read copies of the last slot; theta: reads clamped to the window
  // theta: the theta-major table, whose slots run along the pencil (reads past the window
  // land on the next ring's slots or the overhang: finite values no kept cell depends on)
  const unsigned tbytes = 4u * (unsigned)kTabEntries * (unsigned)f.tstride;
  // A window of at most SL K cells (a short last window, or a whole short pencil) runs
  // NSEG = 64 / SL columns per wavefront pass, one in each run of SL lanes (lanes past the
  // last run idle); every stage indexes cells within its own window (cell_in<SL>), so each
  // column's result is the one-column pass's.  Column group gi = the NSEG columns
  // gi NSEG + seg; wave wv takes groups wv QP .. wv QP + QP - 1; a column past the
  // workgroup's NC reads column NC - 1 and stores nothing.
  auto pencils = [&](auto sl_tag) {
    constexpr int SL = decltype(sl_tag)::value, NSEG = 64 / SL;
    constexpr int NG = (NC + NSEG - 1) / NSEG, QP = (NG + W - 1) / W;
    const int seg = lane / SL, ll = lane % SL;
    const TabLane<K> tx{Buf(TH ? f.tabz : f.tab, tbytes),
                        4u * (unsigned)((TH ? (REAL * l - kNpad + g.k0) + g.sk * (row + kNpad) : s0) + K * ll),
                        f.tstride};
  for (int qq = 0; qq < QP; ++qq) {
    const int gi = wv * QP + qq;
    if (gi >= NG || jb + gi * NSEG >= jmax) break;   // wave-uniform
    const int qv = gi * NSEG + seg;
    const bool mine = seg < NSEG && qv < NC;   // this lane's column is one of the workgroup's
    const int q = mine ? qv : NC - 1;
    // a column past jmax (the last group) computes on column jmax - 1's copies and is never
    // stored
    const int j0 = jb + q < jmax ? jb + q : jmax - 1;
    WinP<K> Wp;
    Wp.rad = tx.get(kTabRad);
    if constexpr (TH) {
      Wp.xa0 = tx.get(kTabZa0);
      Wp.dx0 = tx.get(kTabDz0);
      Wp.dvol0 = tx.get(kTabDvolZ);
      Wp.drb = tx.get(kTabDrbZ);
      Wp.dpb = tx.get(kTabDpbZ);
      Wp.dLb = tx.get(kTabDLbZ);
    } else {
      Wp.xa0 = tx.get(kTabXa0);
      Wp.dx0 = tx.get(kTabDx0);
      Wp.dvol0 = tx.get(kTabDvol0);
      Wp.drb = tx.get(kTabDrb);
      Wp.dpb = tx.get(kTabDpb);
      Wp.dLb = tx.get(kTabDLb);
    }
    Wp.rb = tx.get(kTabRb);
    Wp.pb = tx.get(kTabPb);
    Wp.cs2 = tx.get(kTabCs2);
    Wp.cs = tx.get(kTabCs);
    Wp.Lb = tx.get(kTabLb);
    Wp.rs = tx.get(kTabRotS);
    Wp.rcs = tx.get(kTabRotC);

    if constexpr (G != kGeoRcyl) {
      Wp.sth = tx.get(kTabSth);
      Wp.oms = tx.get(kTabOms);
    }
    if constexpr (ADIA) {
      Wp.qb = tx.get(kTabQb);
      Wp.dqb = tx.get(kTabDqb);
    }
    const float2 tr = f.trig[j0];
    Wp.ts = splat<float, K>(tr.x);
    Wp.tc = splat<float, K>(tr.y);
    const C<float, K> rc = tx.get(kTabRc);
    const C<float, K> f0 = lds_k<K>(&tile[0][q][K * ll]), f1 = lds_k<K>(&tile[1][q][K * ll]),
                      f2 = lds_k<K>(&tile[2][q][K * ll]);
    C<float, K> f3;
    if constexpr (NF == 4) f3 = lds_k<K>(&tile[3][q][K * ll]);
    Wp.r = f0;
#pragma unroll
    for (int h = 0; h < K; ++h) {
      // l - L_b = r_cyl (v_phi - v_bg); spherical radial w = r v_theta
      if constexpr (TH) {
        Wp.u[h] = f3[h];
        Wp.v[h] = f1[h];
        Wp.w[h] = rc[h] * f2[h];
      } else {
        Wp.u[h] = f1[h];
        Wp.v[h] = rc[h] * f2[h];
        if constexpr (G == kGeoRsph) Wp.w[h] = Wp.rad[h] * f3[h];
        if constexpr (ADIA) Wp.q[h] = f3[h];
      }
    }
    Para<float, K> P0;
    if constexpr (TH) {
      P0 = f.uniform_theta ? paraset_uniform<K>() : paraset<float, K, SL>(0, lim, Wp.dx0);
    } else {
      P0.c0 = tx.get(kTabP0);
      P0.c1 = tx.get(kTabP1);
      P0.c2 = tx.get(kTabP2);
      P0.c3 = tx.get(kTabP3);
      P0.c4 = tx.get(kTabP4);
    }
    ppmlr_pert<K, G, (G == kGeoRcyl ? 2 : 3), TabLane<K>, ADIA, SL, BARY>(Wp, tx, P0, lim, dt, 0.0f, pl, prm.plnt > 0,
                                                                     prm.flat != 0, sqrt_gam, 0.0f, false,
                                                                     (float)prm.gam);
    if (VISC == 1 || (VISC == 2 && prm.visc)) {
      if constexpr (G == kGeoRcyl) visc_r_pert<K, SL>(Wp, tx, lim, dt);
      else if constexpr (G == kGeoRsph) visc_r_pert<K, SL, true>(Wp, tx, lim, dt);
      else visc_z_pert<K, SL>(Wp, tx, lim, dt, (float)prm.vis_nu);
    }
    const C<float, K> irc = tx.get(kTabIrc);
    C<float, K> ixc;
    if constexpr (G == kGeoRsph) ixc = tx.get(kTabIxc);
    // angular momenta back to velocities as increments: v_new = v + (l_new - r v) / r with
    // r v formed as on the way in, so a cell whose l did not change keeps v exactly (the
    // product (r v) (1/r) would scale it by the table's fixed r (1/r) - 1 every step)
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const int c = K * ll + h;
      if (mine && c >= kNpad && c < lim - kNpad) {   // only this lane's run touches column q
        tile[0][q][c] = Wp.r[h];
        if constexpr (TH) {
          tile[3][q][c] = Wp.u[h];
          tile[1][q][c] = Wp.v[h];
          const float v0 = tile[2][q][c];
          tile[2][q][c] = v0 + (Wp.w[h] - rc[h] * v0) * irc[h];
        } else {
          tile[1][q][c] = Wp.u[h];
          const float v0 = tile[2][q][c];
          tile[2][q][c] = v0 + (Wp.v[h] - rc[h] * v0) * irc[h];
          if constexpr (G == kGeoRsph) {
            const float w0 = tile[3][q][c];
            tile[3][q][c] = w0 + (Wp.w[h] - Wp.rad[h] * w0) * ixc[h];
          }
          if constexpr (ADIA) tile[3][q][c] = Wp.q[h];
        }
      }
    }
  }
  };
  // workgroup-uniform choice: three windows of 21 lanes (K = 4: a 64-cell theta pencil with
  // its halos, 76 cells), two of 32, or one
  bool three = false;
  if constexpr (K == 4) three = f.pack_tails && lim <= 21 * K;
  if (three) {
    if constexpr (K == 4) pencils(std::integral_constant<int, 21>{});
  } else if (f.pack_tails && lim <= 32 * K) {
    pencils(std::integral_constant<int, 32>{});
  } else {
    pencils(std::integral_constant<int, 64>{});
  }
  __syncthreads();
  float m = 0.0f;
  const double FrRot = ss->FrRot;
  Buf fout[NF];
#pragma unroll
  for (int fi = 0; fi < NF; ++fi) fout[fi] = Buf(f.f[src ^ 1][fi], fbytes);
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int q = qs, c = cs + CSTEP * it;
    if (jl < jmax && c >= kNpad && c < lim - kNpad) {
      const int sl = s0 + step * c;
      const int j = wrapj(jl - T.srot[c], jmax);
      const unsigned o = 4u * (unsigned)(sl * jmax + j);
#pragma unroll
      for (int fi = 0; fi < NF; ++fi) fout[fi].put(tile[fi][q][c], o);
      if constexpr (CFL) {
        // the next step's CFL term of this cell (timestep.cu:12-108), isothermal c_s
        const float* T = f.tab + sl;
        const long ts = f.tstride;
        const float csnd = sqrt_gam * T[kTabCs * ts];
        const float cfl_off = (float)(f.vbg[sl] - g.rot_v[sl] - FrRot * g.radc[sl]);
        const float t0 = (fabsf(tile[1][q][c]) + csnd) / T[kTabDx0 * ts];
        const float t1 = (fabsf(tile[2][q][c] + cfl_off) + csnd) / (f.dy[j] * T[kTabRc * ts]);
        const float t2 = (fabsf(tile[NF - 1][q][c]) + csnd) / (T[kTabDz0 * ts] * T[kTabRad * ts]);
        m = fmaxf(m, fmaxf(t0, fmaxf(t1, t2)));
      }
    }
  }
  if constexpr (CFL) {
    m = wave_maxf(m);
    if (lane == 0) T.red[wv] = m;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int w = 1; w < W; ++w) m = fmaxf(m, T.red[w]);
  }
  return m;
}

// one disk: one workgroup per item, XCD-contiguous
template <int K, int W, int Q, int VISC, bool ADIA, bool BARY>
__global__ __launch_bounds__(64 * W) PN_WAVES((pencil_waves<K, W, Q, ADIA ? 4 : 3>())) void k_sweep_x_f32(DevGrid g, DevF32 f, DevParams prm,
                                                        const StepScalars* __restrict__ ss, int src) {
  __shared__ XTile<K, W, Q, ADIA ? 4 : 3> T;
  const int ngrp = (g.jmax + W * Q - 1) / (W * Q);
  const int bid = xcd_swizzle(blockIdx.x, gridDim.x);
  const int grp = bid % ngrp, l = bid / ngrp;
  if (l >= f.nwin_x) return;   // workgroup-uniform, before any barrier
  pencil_item<K, W, Q, kGeoRcyl, VISC, false, ADIA, BARY>(g, f, prm, ss, src, grp, l, 0, T);
}

// 3-D: the spherical radial sweep (item = theta row k, radial window, column group) and the
// theta sweep (item = ring i, theta window, column group; with do_cfl it is the step's last
// sweep and writes one CFL partial per workgroup)
template <int K, int W, int Q, int G, int VISC, bool BARY>
__global__ __launch_bounds__(64 * W) PN_WAVES((pencil_waves<K, W, Q, 4>())) void k_sweep_pencil3_f32(DevGrid g, DevF32 f, DevParams prm,
                                                              const StepScalars* __restrict__ ss, double* cflp,
                                                              int src, int do_cfl) {
  __shared__ XTile<K, W, Q, 4> T;
  const int ngrp = (g.jmax + W * Q - 1) / (W * Q);
  const int nwin = (G == kGeoTheta) ? f.nwin_z : f.nwin_x;
  const int nrow = (G == kGeoTheta) ? g.imax : g.kmax;
  const int bid = xcd_swizzle(blockIdx.x, gridDim.x);
  const int grp = bid % ngrp, rest = bid / ngrp;
  const int l = rest % nwin, row = rest / nwin;
  if (row >= nrow) return;   // workgroup-uniform, before any barrier
  if (do_cfl) {
    const float m = pencil_item<K, W, Q, G, VISC, true, false, BARY>(g, f, prm, ss, src, grp, l, row, T);
    if (threadIdx.x == 0) cflp[blockIdx.x] = (double)m;
  } else {
    pencil_item<K, W, Q, G, VISC, false, false, BARY>(g, f, prm, ss, src, grp, l, row, T);
  }
}

// the disk owning work item it: base[d] <= it < base[d + 1] (ndisk small; base in device memory)
__device__ __forceinline__ int disk_of(const int* __restrict__ base, int ndisk, int it) {
  int lo = 0, hi = ndisk - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= it) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// many disks, persistent: workgroup b runs items t = b, b + G, ... (G = gridDim, a multiple
// of 8), each mapped through the XCD permutation so that one XCD walks a contiguous run of
// items (neighbouring column groups share L2 lines).  base: ndisk + 1 item offsets.
// The batch's radial kernel also keeps its item's disk lookup in scalar registers: asked for
// five waves per SIMD it spilled 24 VGPRs to scratch and 60 SGPRs to VGPR lanes (each a VALU
// lane write or read), at four it spills nothing and ran 7 % faster on the 48-disk sweep
// (0.428 -> 0.397 ms per launch, same box; six waves: 0.428).
template <int W>
constexpr int batch_pencil_waves() {
  return W == 4 ? 4 : 1;
}

// order: nslots (disk, item) pairs, the item of slot t for the workgroup that takes it
// (slot t runs on XCD t % 8; the host lays each XCD's share out by decreasing cost, disk -1
// marks an empty slot)
template <int K, int W, int Q, bool BARY, bool ADIA>
__global__ __launch_bounds__(64 * W) PN_WAVES(ADIA ? 1 : batch_pencil_waves<W>()) void k_sweep_x_f32_batch(const DiskDev* __restrict__ dd,
                                                              const DevF32* __restrict__ dfs,
                                                              const int* __restrict__ order, int nslots, int src) {
  __shared__ XTile<K, W, Q, ADIA ? 4 : 3> T;
  for (int t = blockIdx.x; t < nslots; t += gridDim.x) {
    const int d = order[2 * t];
    if (d < 0) continue;   // workgroup-uniform
    const DevGrid g = dd[d].g;
    const int ngrp = (g.jmax + W * Q - 1) / (W * Q);
    const int local = order[2 * t + 1];
    pencil_item<K, W, Q, kGeoRcyl, 2, false, ADIA, BARY>(g, dfs[d], dd[d].p, dd[d].ss, src, local % ngrp, local / ngrp,
                                                        0, T);
  }
}

// ---------------------------------------------------------------------------------------
// Azimuthal sweep of window l of interior ring n (3-D: ring n % imax of theta row n / imax)
// by one wavefront; returns the wave's maximum CFL term of the cells it stored (2-D).  The
// sweep's radius is r_cyl; in 3-D it also carries v_theta and the planet's spherical forms.
template <int K, int VISC, int NDIM, bool ADIA = false, bool BARY = false>
__device__ __forceinline__ float y_pencil(const DevGrid& g, const DevF32& f, const DevParams& prm,
                                          const StepScalars* __restrict__ ss, int src, int n, int l) {
  constexpr int REAL = 64 * K - 2 * kNpad;
  const int lane = threadIdx.x & 63;
  const int s = (NDIM == 3) ? slot_of(g, n % g.imax, n / g.imax) : slot_of(g, n, 0);
  const int jmax = g.jmax;
  const float dt = (float)ss->dt;
  const double FrRot = ss->FrRot;
  const PlanetF pl = planet_f<BARY>(ss, NDIM == 3, prm.bary);
  const int lim = win_len_k<K>(jmax, f.nwin_y, l);
  const int rot = g.rot_j[s];
  const double rcd = g.radc[s];
  const double vbg = f.vbg[s];
  // grid motion of the shifted sweep: (v_bg - res_v)/r - FrRot; the v_phi offset the CFL
  // sees: v_bg - rot_v - FrRot r (timestep.cu:62)
  const float wres = (float)((vbg - g.res_v[s]) / rcd - FrRot);
  const float cfl_off = (float)(vbg - g.rot_v[s] - FrRot * rcd);
  const float* __restrict__ T = f.tab + s;
  const long ts = f.tstride;
  const float rc = T[kTabRc * ts], dxr = T[kTabDx0 * ts], rb = T[kTabRb * ts], pb = T[kTabPb * ts],
              cs2 = T[kTabCs2 * ts], cs = T[kTabCs * ts], vrot = T[kTabVrot * ts], rps = T[kTabRotPhS * ts],
              rpc = T[kTabRotPhC * ts];
  float rsph = 0.0f, sth = 1.0f, oms = 0.0f, qb = 0.0f;
  if constexpr (ADIA) qb = T[kTabQb * ts];
  if constexpr (NDIM == 3) {
    rsph = T[kTabRad * ts];
    sth = T[kTabSth * ts];
    oms = T[kTabOms * ts];
  }
  // the ring's fields and the column tables as buffer views (byte offsets 4 j)
  const long base = (long)s * jmax;
  const unsigned rb4 = 4u * (unsigned)jmax;
  const Buf bfr(f.f[src][0] + base, rb4), bfu(f.f[src][1] + base, rb4), bfv(f.f[src][2] + base, rb4);
  const Buf btrig(f.trig, 2u * rb4), bdy(f.dy, rb4), byv(f.yvol, rb4);

  WinP<K> W;
  unsigned jd[K];   // byte offset of the cell's data column
  C<float, K> dyc;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const int c = K * lane + h;
    // lab column, then data column: each needs at most one wrap (rot_j in [0, jmax))
    int j0 = REAL * l - kNpad + (c < lim ? c : lim - 1);
    j0 += (j0 < 0) ? jmax : 0;
    j0 -= (j0 >= jmax) ? jmax : 0;
    int j = j0 - rot;
    j += (j < 0) ? jmax : 0;
    const unsigned o0 = 4u * (unsigned)j0, o = 4u * (unsigned)j;
    jd[h] = o;
    const float2 tr = btrig.f2(2u * o0);
    W.ts[h] = tr.x;
    W.tc[h] = tr.y;
    W.dx0[h] = bdy.f(o0);
    W.dvol0[h] = byv.f(o0) * rc;
    W.xa0[h] = 0.0f;
    W.rad[h] = rc;
    W.vrot[h] = vrot;
    W.rs[h] = rps;
    W.rcs[h] = rpc;
    W.rb[h] = rb;
    W.drb[h] = 0.0f;
    W.pb[h] = pb;
    W.dpb[h] = 0.0f;
    W.cs2[h] = cs2;
    W.cs[h] = cs;
    W.Lb[h] = 0.0f;
    W.dLb[h] = 0.0f;
    W.sth[h] = sth;
    W.oms[h] = oms;
    W.qb[h] = qb;
    W.dqb[h] = 0.0f;
    dyc[h] = bdy.f(o);
    W.r[h] = bfr.f(o);
    W.u[h] = bfv.f(o);   // v_phi - v_bg
    W.v[h] = bfu.f(o);   // v_r
    if constexpr (NDIM == 3) W.w[h] = Buf(f.f[src][3] + base, rb4).f(o);
    if constexpr (ADIA) W.q[h] = Buf(f.f[src][3] + base, rb4).f(o);
  }
  const float sqrt_gam = sqrtf((float)prm.gam);
  const Para<float, K> P0 = f.uniform_phi ? paraset_uniform<K>() : paraset(0, lim, W.dx0);
  ppmlr_pert<K, kGeoPhi, NDIM, TabLane<K>, ADIA, 64, BARY>(W, TabLane<K>{}, P0, lim, dt, wres, pl, prm.plnt > 0,
                                                           prm.flat != 0, sqrt_gam, rsph, f.uniform_phi != 0,
                                                           (float)prm.gam);
  if (VISC == 1 || (VISC == 2 && prm.visc)) visc_p_pert<K, NDIM == 3>(W, lim, dt, T[kTabNu * ts], rc);

  const Buf dr(f.f[src ^ 1][0] + base, rb4), du(f.f[src ^ 1][1] + base, rb4), dv(f.f[src ^ 1][2] + base, rb4);
  const float csnd = sqrt_gam * cs;
  const float idx = 1.0f / dxr;
  float m = 0.0f;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const int c = K * lane + h;
    if (c >= kNpad && c < lim - kNpad) {
      const unsigned o = jd[h];
      dr.put(W.r[h], o);
      du.put(W.v[h], o);
      dv.put(W.u[h], o);
      if constexpr (NDIM == 3) Buf(f.f[src ^ 1][3] + base, rb4).put(W.w[h], o);
      if constexpr (ADIA) Buf(f.f[src ^ 1][3] + base, rb4).put(W.q[h], o);
      // the next step's CFL term of this cell (timestep.cu:48-56); EOS 3: c_s^2 = gam q
      float cn = csnd;
      if constexpr (ADIA) cn = sqrtf((float)prm.gam * (qb + W.q[h]));
      const float t0 = (fabsf(W.v[h]) + cn) * idx;
      const float t1 = (fabsf(W.u[h] + cfl_off) + cn) / (dyc[h] * rc);
      m = fmaxf(m, fmaxf(t0, t1));
    }
  }
  return wave_maxf(m);
}

// one wavefront per (ring, window); with do_cfl each workgroup writes the maximum CFL term
// of its waves to cflp[workgroup] (2-D)
template <int K, int WPB, int VISC, int NDIM, bool ADIA, bool BARY>
__global__ __launch_bounds__(64 * WPB) PN_WAVES((K == 2 && NDIM == 2 && !ADIA) ? 8 : 1) void k_sweep_y_f32(DevGrid g, DevF32 f, DevParams prm,
                                                          const StepScalars* __restrict__ ss, double* cflp, int src,
                                                          int do_cfl) {
  __shared__ float shm[WPB];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform, in a scalar register
  const int gw = xcd_swizzle(blockIdx.x, gridDim.x) * WPB + wv;
  const int l = gw % f.nwin_y;
  const int n = gw / f.nwin_y;
  float m = 0.0f;
  if (n < g.imax * g.kmax) m = y_pencil<K, VISC, NDIM, ADIA, BARY>(g, f, prm, ss, src, n, l);   // wave-uniform
  if (do_cfl) {   // kernel-uniform: every wave reaches the barrier
    if ((threadIdx.x & 63) == 0) shm[wv] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < WPB; ++w) m = fmaxf(m, shm[w]);
      cflp[blockIdx.x] = (double)m;
    }
  }
}

// many disks, persistent (as k_sweep_x_f32_batch): item = one workgroup's WPB waves of one
// disk; each item writes its CFL partial to its disk's cflp[item - base[d]]
template <int K, int WPB, bool BARY, bool ADIA>
__global__ __launch_bounds__(64 * WPB) PN_WAVES((K == 2 && !ADIA) ? 6 : 1) void k_sweep_y_f32_batch(const DiskDev* __restrict__ dd,
                                                              const DevF32* __restrict__ dfs,
                                                              const int* __restrict__ base, int ndisk, int src) {
  __shared__ float shm[WPB];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform, in a scalar register
  const int nitems = base[ndisk];
  for (int t = blockIdx.x; t < nitems; t += gridDim.x) {
    const int it = xcd_swizzle(t, nitems);
    const int d = disk_of(base, ndisk, it);
    const DevGrid g = dd[d].g;
    const DevF32& f = dfs[d];
    const int local = it - base[d];
    const int gw = local * WPB + wv;
    const int l = gw % f.nwin_y;
    const int n = gw / f.nwin_y;
    float m = 0.0f;
    if (n < g.imax) m = y_pencil<K, 2, 2, ADIA, BARY>(g, f, dd[d].p, dd[d].ss, src, n, l);   // wave-uniform
    __syncthreads();   // the previous item's partial has been read
    if ((threadIdx.x & 63) == 0) shm[wv] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < WPB; ++w) m = fmaxf(m, shm[w]);
      dd[d].cflp[local] = (double)m;
    }
  }
}

__device__ __forceinline__ float block_max256(float m, float* sh) {
  m = wave_maxf(m);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  return fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
}

// CFL from the stored FP32 fields, one partial per block
__global__ __launch_bounds__(256) void k_cfl_f32(DevGrid g, DevF32 f, DevParams prm, const StepScalars* ss,
                                                 double* cflp, int src) {
  __shared__ float sh[4];
  const long ncell = (long)g.imax * g.kmax * g.jmax;
  const double FrRot = ss->FrRot;
  const float sqrt_gam = sqrtf((float)prm.gam);
  float m = 0.0f;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += (long)gridDim.x * blockDim.x) {
    const int j = (int)(t % g.jmax);
    const int n = (int)(t / g.jmax);
    const int s = slot_of(g, n % g.imax, n / g.imax);
    const long a = (long)s * g.jmax + j;
    const float* T = f.tab + s;
    const long ts = f.tstride;
    // EOS 3: c_s^2 = gam q (2-D; field 3 carries q - q_b)
    const float csnd = (prm.eos == 3) ? sqrtf((float)prm.gam * (T[kTabQb * ts] + f.f[src][3][a]))
                                      : sqrt_gam * T[kTabCs * ts];
    const float cfl_off = (float)(f.vbg[s] - g.rot_v[s] - FrRot * g.radc[s]);
    const float t0 = (fabsf(f.f[src][1][a]) + csnd) / T[kTabDx0 * ts];
    const float t1 = (fabsf(f.f[src][2][a] + cfl_off) + csnd) / (f.dy[j] * T[kTabRc * ts]);
    m = fmaxf(m, fmaxf(t0, t1));
    if (g.ndim == 3) m = fmaxf(m, (fabsf(f.f[src][3][a]) + csnd) / (T[kTabDz0 * ts] * T[kTabRad * ts]));
  }
  m = block_max256(m, sh);
  if (threadIdx.x == 0) cflp[blockIdx.x] = (double)m;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// mass (mode 0) and the planet-force integrands (modes 8, 9) of device_func.cu:49-198 on
// the FP32 state, in FP64 with the FP64 background added back
__global__ __launch_bounds__(256) void k_reduce_f32(DevGrid g, DevF32 f, DevParams prm,
                                                    const StepScalars* __restrict__ ss, int src, int mode,
                                                    double exclude, double* partials) {
  __shared__ double sh[4];
  const long ncell = (long)g.imax * g.kmax * g.jmax;
  const Planet pl{ss->planet.m, ss->planet.x, ss->planet.y, ss->planet.vx, ss->planet.vy, ss->planet.rs, prm.bary};
  double acc = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += (long)gridDim.x * blockDim.x) {
    const int j = (int)(t % g.jmax);
    const int n = (int)(t / g.jmax);
    const int s = slot_of(g, n % g.imax, n / g.imax);
    int j0 = j + g.rot_j[s];
    if (j0 >= g.jmax) j0 -= g.jmax;
    const double r = f.rbg[s] + (double)f.f[src][0][(long)s * g.jmax + j];
    double tmp = 0.0;
    if (mode == 0 || prm.plnt == 0) {
      tmp = r * g.xvol[s] * g.yvol[j0] * g.zvol[s];
    } else if (g.ndim == 2 && (mode == 8 || mode == 9)) {
      const double rad = g.xc[s], phi = g.yc[j0];
      const double dd = (rad - pl.x) * (rad - pl.x) + rad * rad * (phi - pl.y) * (phi - pl.y);
      if (dd > exclude * exclude) {
        const double bg = pow(g.xc[s], -prm.p_alpha);
        if (mode == 8) tmp = -grav_rad_cyl<double>(pl, rad, phi, 0.0, 0.0) * (r - bg) * g.xvol[s] * g.yvol[j0];
        else tmp = -grav_azi_cyl<double>(pl, rad, phi, 0.0, 0.0) * (rad * (r - bg) * g.xvol[s] * g.yvol[j0]);
      }
    }
    acc += tmp;
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// Damping zones (kill_wave.cu:3-43) on the FP32 state: the rings within sc_h/2 of xmin or xmax
// relax towards their background record by the fraction k = dt r^-1.5 (distance into the zone) /
// (sc_h/2), as the FP64 path does.  The background record is the FP32 state's own background,
// so Sigma - Sigma_b, v_phi - v_bg and v_theta scale by 1 - k and v_r relaxes towards the
// record's v_r; EOS 3 relaxes p and Sigma and carries q = p / Sigma - q_b again, in FP64.
// Thread: one cell of an interior ring.
// set_boundx (boundary.cu:86-206) on the FP32 state of buffer src: the copies of the FP64
// path's k_fill_ghosts, of deviations.  A ghost ring a code copies into carries its source
// ring's backgrounds (set on load), so the copied deviation is the copied value; code 0
// negates v_r, code 4 copies Sigma and P (and v_r on the left) and leaves v_phi and v_theta
// fixed.  Fields: 0 Sigma, 1 v_r, 2 v_phi, 3 v_theta (3-D) or q = P / Sigma (EOS 3).
// Thread: one lab column of one of the 12 ghost rings of a theta row.
__global__ void k_fill_ghosts_f32(DevGrid g, DevF32 f, DevParams prm, int src) {
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
  f.f[src][0][ag] = f.f[src][0][as];
  if (code != 4 || left) f.f[src][1][ag] = (code == 0) ? -f.f[src][1][as] : f.f[src][1][as];
  if (code != 4) f.f[src][2][ag] = f.f[src][2][as];
  if ((g.ndim == 3 && code != 4) || prm.eos == 3) f.f[src][3][ag] = f.f[src][3][as];
}

__global__ void k_damp_f32(DevGrid g, DevF32 f, DevParams prm, const StepScalars* __restrict__ ss, int src) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= g.jmax) return;
  const int n = blockIdx.y;   // ring i + imax k
  const int s = slot_of(g, n % g.imax, n / g.imax);
  const double d_kill = 0.5 * prm.sc_h;
  const double inner = prm.xmin + d_kill, outer = prm.xmax - d_kill;
  const double r = g.xc[s];
  double k;
  if (r < inner) k = (inner - r) / d_kill;
  else if (r > outer) k = (r - outer) / d_kill;
  else return;
  k *= ss->dt * pow(r, -1.5);
  const pn_symdisk V = g.val[n];
  const long a = (long)s * g.jmax + j;
  float* const F0 = f.f[src][0];
  float* const F1 = f.f[src][1];
  float* const F2 = f.f[src][2];
  const double dr = F0[a];
  F0[a] = (float)((1.0 - k) * dr);
  F1[a] = (float)((double)F1[a] + k * (V.u - (double)F1[a]));
  F2[a] = (float)((1.0 - k) * (double)F2[a]);
  if (g.ndim == 3 || prm.eos == 3) {   // the fourth field exists
    float* const F3 = f.f[src][3];
    if (g.ndim == 3) {
      F3[a] = (float)((1.0 - k) * (double)F3[a]);
    } else {
      const double qb = V.p / V.r, rho = V.r + dr;
      const double p = rho * (qb + (double)F3[a]);
      const double p1 = p + k * (V.p - p), rho1 = V.r + (1.0 - k) * dr;
      F3[a] = (float)(p1 / rho1 - qb);
    }
  }
}

__global__ __launch_bounds__(256) void k_totals_f32(DevGrid g, DevF32 f, int src, double* partials) {
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
    const double dm = (f.rbg[s] + (double)f.f[src][0][a]) * g.xvol[s] * g.yvol[j0] * g.zvol[s];
    const double dl = dm * (f.vbg[s] + (double)f.f[src][2][a]) * g.radc[s];
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

// Azimuthal mean of a field of every interior ring on the FP32 state (the FP64 path's
// k_ring_means, same dphi weighting and order): the deviations from the ring's background are
// summed in FP64 and the background added back once, so the mean of a ring that departs little
// from its background keeps every digit the deviations hold.  field: 0 Sigma, 1 P (EOS 0:
// c_s^2 Sigma; EOS 3: Sigma q), 2 v_r, 3 v_phi, 4 v_theta (3-D; 0 in 2-D).
__global__ __launch_bounds__(256) void k_ring_means_f32(DevGrid g, DevF32 f, int eos, int src, int field,
                                                        double* __restrict__ out) {
  __shared__ double sh[8];
  const int n = blockIdx.x;
  const int s = slot_of(g, n % g.imax, n / g.imax);
  const long base = (long)s * g.jmax;
  const int rot = g.rot_j[s];
  const double rb = f.rbg[s], cs2 = g.cs2[s];
  const double qb = (eos == 3) ? g.val[n].p / g.val[n].r : 0.0;
  const float* __restrict__ F0 = f.f[src][0] + base;
  const float* __restrict__ F1 = f.f[src][1] + base;
  const float* __restrict__ F2 = f.f[src][2] + base;
  const float* __restrict__ F3 = (g.ndim == 3 || eos == 3) ? f.f[src][3] + base : nullptr;
  double a = 0.0, w = 0.0;
  for (int j = threadIdx.x; j < g.jmax; j += 256) {
    int j0 = j + rot;
    if (j0 >= g.jmax) j0 -= g.jmax;
    const double d = g.dy[j0];
    double x = 0.0;
    if (field == 0) {
      x = F0[j];
    } else if (field == 1) {
      const double dr = F0[j];
      x = (eos == 3) ? rb * (double)F3[j] + qb * dr + dr * (double)F3[j] : cs2 * dr;
    } else if (field == 2) {
      x = F1[j];
    } else if (field == 3) {
      x = F2[j];
    } else if (g.ndim == 3) {
      x = F3[j];
    }
    a += x * d;
    w += d;
  }
  a = wave_sum(a);
  w = wave_sum(w);
  if ((threadIdx.x & 63) == 0) {
    sh[threadIdx.x >> 6] = a;
    sh[4 + (threadIdx.x >> 6)] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double bg = (field == 0) ? rb : (field == 1) ? ((eos == 3) ? rb * qb : cs2 * rb) : (field == 3) ? f.vbg[s] : 0.0;
    out[n] = bg + ((sh[0] + sh[1]) + (sh[2] + sh[3])) / ((sh[4] + sh[5]) + (sh[6] + sh[7]));
  }
}

int red_blocks(const DevGrid& g) {
  long ncell = (long)g.imax * g.kmax * g.jmax;
  long b = (ncell + 255) / 256;
  if (b > 2048) b = 2048;
  return (int)(b < 1 ? 1 : b);
}

template <int K, int W, int Q>
hipError_t sweep_x_k(const LaunchCtx& c, const DevF32& f0, int src) {
  DevF32 f = f0;
  f.nwin_x = f32_windows(c.g.imax, K);
  // the last window reads table entries up to 64 K - 1 slots past its first slot
  if (f.tstride < c.g.sr * c.g.sk + 64 * K) return hipErrorInvalidValue;
  const long ngrp = (c.g.jmax + W * Q - 1) / (W * Q);
  dim3 grid((unsigned)(ngrp * f.nwin_x)), block(64 * W);
  // the same body as the batched kernels (viscosity switched by the parameters), so a disk
  // computes the same bits alone or in a batch
  // the barycentric frame's star terms are compiled into kernels of their own (BARY), so the
  // others run without them
  auto go = [&](auto bary) {
    constexpr bool B = decltype(bary)::value;
    // EOS 3: a viscous disk runs its own instantiation (the viscosity updates velocities only)
    if (c.p.eos == 3 && c.p.visc) k_sweep_x_f32<K, W, Q, 1, true, B><<<grid, block, 0, c.stream>>>(c.g, f, c.p, c.ss, src);
    else if (c.p.eos == 3) k_sweep_x_f32<K, W, Q, 0, true, B><<<grid, block, 0, c.stream>>>(c.g, f, c.p, c.ss, src);
    else k_sweep_x_f32<K, W, Q, 2, false, B><<<grid, block, 0, c.stream>>>(c.g, f, c.p, c.ss, src);
  };
  if (c.p.bary) go(std::true_type{});
  else go(std::false_type{});
  return hipGetLastError();
}

template <int K, int WPB>
hipError_t sweep_y_k(const LaunchCtx& c, const DevF32& f0, int src, int do_cfl) {
  DevF32 f = f0;
  f.nwin_y = f32_windows(c.g.jmax, K);
  const long waves = (long)f.nwin_y * c.g.imax * c.g.kmax;
  const long nblk = (waves + WPB - 1) / WPB;
  if (do_cfl && (nblk > c.ncflp || c.g.ndim == 3)) return hipErrorInvalidValue;
  dim3 grid((unsigned)nblk), block(64 * WPB);
  auto go = [&](auto bary) {
    constexpr bool B = decltype(bary)::value;
    if (c.g.ndim == 3 && c.p.visc)
      k_sweep_y_f32<K, WPB, 1, 3, false, B><<<grid, block, 0, c.stream>>>(c.g, f, c.p, c.ss, c.cflp, src, 0);
    else if (c.g.ndim == 3)
      k_sweep_y_f32<K, WPB, 0, 3, false, B><<<grid, block, 0, c.stream>>>(c.g, f, c.p, c.ss, c.cflp, src, 0);
    else if (c.p.eos == 3 && c.p.visc)
      k_sweep_y_f32<K, WPB, 1, 2, true, B><<<grid, block, 0, c.stream>>>(c.g, f, c.p, c.ss, c.cflp, src, do_cfl);
    else if (c.p.eos == 3)
      k_sweep_y_f32<K, WPB, 0, 2, true, B><<<grid, block, 0, c.stream>>>(c.g, f, c.p, c.ss, c.cflp, src, do_cfl);
    else k_sweep_y_f32<K, WPB, 2, 2, false, B><<<grid, block, 0, c.stream>>>(c.g, f, c.p, c.ss, c.cflp, src, do_cfl);
  };
  if (c.p.bary) go(std::true_type{});
  else go(std::false_type{});
  return hipGetLastError();
}

// 3-D radial (G = kGeoRsph) and theta (kGeoTheta) sweeps; items: column group x window x row
template <int K, int W, int Q, int G>
hipError_t sweep_pencil3_k(const LaunchCtx& c, const DevF32& f0, int src, int do_cfl) {
  DevF32 f = f0;
  const bool th = (G == kGeoTheta);
  if (th) f.nwin_z = f32_windows(c.g.kmax, K);
  else f.nwin_x = f32_windows(c.g.imax, K);
  if (c.g.ndim != 3 || f.f[0][3] == nullptr || f.tabz == nullptr) return hipErrorInvalidValue;
  // theta windows read the theta-major table up to 64 K - 1 slots past their first slot
  if (th && f.tstride < c.g.sr * c.g.sk + 64 * K) return hipErrorInvalidValue;
  // radial windows read table entries up to 64 K - 1 slots past their first slot
  if (!th && f.tstride < c.g.sr * c.g.sk + 64 * K) return hipErrorInvalidValue;
  const long ngrp = (c.g.jmax + W * Q - 1) / (W * Q);
  const long nblk = ngrp * (th ? (long)f.nwin_z * c.g.imax : (long)f.nwin_x * c.g.kmax);
  if (do_cfl && (!th || nblk > c.ncflp)) return hipErrorInvalidValue;
  dim3 grid((unsigned)nblk), block(64 * W);
  // a viscous disk runs its own instantiation, so the inviscid one keeps its registers
  auto go = [&](auto bary) {
    constexpr bool B = decltype(bary)::value;
    if (c.p.visc)
      k_sweep_pencil3_f32<K, W, Q, G, 1, B><<<grid, block, 0, c.stream>>>(c.g, f, c.p, c.ss, c.cflp, src, do_cfl);
    else k_sweep_pencil3_f32<K, W, Q, G, 0, B><<<grid, block, 0, c.stream>>>(c.g, f, c.p, c.ss, c.cflp, src, do_cfl);
  };
  if (c.p.bary) go(std::true_type{});
  else go(std::false_type{});
  return hipGetLastError();
}

}  // namespace

int f32_windows(int len, int K) {
  const int real = 64 * K - 2 * kNpad;
  return (len + real - 1) / real;
}

const int kF32VariantsX[kNumF32Variants][3] = {{2, 4, 2}, {2, 8, 1}, {2, 16, 1}, {2, 4, 4}};
const int kF32VariantsZ[kNumF32VariantsZ][3] = {{2, 4, 2}, {2, 8, 1}, {2, 16, 1}, {2, 4, 4}, {3, 4, 2}};
const int kF32VariantsY[kNumF32Variants][2] = {{2, 2}, {2, 4}, {4, 2}, {4, 4}};

int f32_cfl_partials_z(const DevGrid& g, int variant) {
  const int nc = kF32VariantsZ[variant][1] * kF32VariantsZ[variant][2];
  return ((g.jmax + nc - 1) / nc) * f32_windows(g.kmax, kF32VariantsZ[variant][0]) * g.imax;
}

int f32_cfl_partials_y(const DevGrid& g, int variant) {
  const int waves = f32_windows(g.jmax, kF32VariantsY[variant][0]) * g.imax * g.kmax;
  const int wpb = kF32VariantsY[variant][1];
  return (waves + wpb - 1) / wpb;
}

hipError_t f32_prep(const LaunchCtx& c, const DevF32& f) {
  const int nslot = c.g.sr * c.g.sk;
  const int n = nslot > c.g.jmax ? nslot : c.g.jmax;
  k_prep<<<(n + 255) / 256, 256, 0, c.stream>>>(c.g, f, c.ss, nslot);
  return hipGetLastError();
}

hipError_t f32_sweep_x(const LaunchCtx& c, const DevF32& f, int variant, int src) {
  switch (variant) {
    case 0: return sweep_x_k<2, 4, 2>(c, f, src);
    case 1: return sweep_x_k<2, 8, 1>(c, f, src);
    case 2: return sweep_x_k<2, 16, 1>(c, f, src);
    case 3: return sweep_x_k<2, 4, 4>(c, f, src);
  }
  return hipErrorInvalidValue;
}

hipError_t f32_sweep_x3(const LaunchCtx& c, const DevF32& f, int variant, int src) {
  switch (variant) {
    case 0: return sweep_pencil3_k<2, 4, 2, kGeoRsph>(c, f, src, 0);
    case 1: return sweep_pencil3_k<2, 8, 1, kGeoRsph>(c, f, src, 0);
    case 2: return sweep_pencil3_k<2, 16, 1, kGeoRsph>(c, f, src, 0);
    case 3: return sweep_pencil3_k<2, 4, 4, kGeoRsph>(c, f, src, 0);
  }
  return hipErrorInvalidValue;
}

hipError_t f32_sweep_z(const LaunchCtx& c, const DevF32& f, int variant, int src, int do_cfl) {
  switch (variant) {
    case 0: return sweep_pencil3_k<2, 4, 2, kGeoTheta>(c, f, src, do_cfl);
    case 1: return sweep_pencil3_k<2, 8, 1, kGeoTheta>(c, f, src, do_cfl);
    case 2: return sweep_pencil3_k<2, 16, 1, kGeoTheta>(c, f, src, do_cfl);
    case 3: return sweep_pencil3_k<2, 4, 4, kGeoTheta>(c, f, src, do_cfl);
    case 4: return sweep_pencil3_k<3, 4, 2, kGeoTheta>(c, f, src, do_cfl);
  }
  return hipErrorInvalidValue;
}

hipError_t f32_sweep_y(const LaunchCtx& c, const DevF32& f, int variant, int src, int do_cfl) {
  switch (variant) {
    case 0: return sweep_y_k<2, 2>(c, f, src, do_cfl);
    case 1: return sweep_y_k<2, 4>(c, f, src, do_cfl);
    case 2: return sweep_y_k<4, 2>(c, f, src, do_cfl);
    case 3: return sweep_y_k<4, 4>(c, f, src, do_cfl);
  }
  return hipErrorInvalidValue;
}

namespace {

// resident workgroups of kernel k on the whole device, rounded down to a multiple of 8 XCDs
template <class Kern>
int persistent_grid(Kern k, int block) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, block, 0) != hipSuccess) return 0;
  const int g = cus * (per_cu > 0 ? per_cu : 1);
  return g >= 8 ? g / 8 * 8 : 8;
}

template <int K, int W, int Q>
hipError_t sweep_x_batch_k(hipStream_t s, const DiskDev* dd, const DevF32* dfs, const int* order, int nslots,
                           int src, int grid, bool bary, bool adia) {
  if (grid < 8 || grid % 8 || nslots % 8) return hipErrorInvalidValue;
  if (bary && adia) k_sweep_x_f32_batch<K, W, Q, true, true><<<grid, 64 * W, 0, s>>>(dd, dfs, order, nslots, src);
  else if (bary) k_sweep_x_f32_batch<K, W, Q, true, false><<<grid, 64 * W, 0, s>>>(dd, dfs, order, nslots, src);
  else if (adia) k_sweep_x_f32_batch<K, W, Q, false, true><<<grid, 64 * W, 0, s>>>(dd, dfs, order, nslots, src);
  else k_sweep_x_f32_batch<K, W, Q, false, false><<<grid, 64 * W, 0, s>>>(dd, dfs, order, nslots, src);
  return hipGetLastError();
}

template <int K, int WPB>
hipError_t sweep_y_batch_k(hipStream_t s, const DiskDev* dd, const DevF32* dfs, const int* base, int ndisk,
                           int src, int grid, bool bary, bool adia) {
  if (grid < 8 || grid % 8) return hipErrorInvalidValue;
  if (bary && adia) k_sweep_y_f32_batch<K, WPB, true, true><<<grid, 64 * WPB, 0, s>>>(dd, dfs, base, ndisk, src);
  else if (bary) k_sweep_y_f32_batch<K, WPB, true, false><<<grid, 64 * WPB, 0, s>>>(dd, dfs, base, ndisk, src);
  else if (adia) k_sweep_y_f32_batch<K, WPB, false, true><<<grid, 64 * WPB, 0, s>>>(dd, dfs, base, ndisk, src);
  else k_sweep_y_f32_batch<K, WPB, false, false><<<grid, 64 * WPB, 0, s>>>(dd, dfs, base, ndisk, src);
  return hipGetLastError();
}

}  // namespace

namespace {
template <bool B, bool A>
int batch_grid(int axis, int variant) {
  if (axis == 0) {
    switch (variant) {
      case 0: return persistent_grid(k_sweep_x_f32_batch<2, 4, 2, B, A>, 256);
      case 1: return persistent_grid(k_sweep_x_f32_batch<2, 8, 1, B, A>, 512);
      case 2: return persistent_grid(k_sweep_x_f32_batch<2, 16, 1, B, A>, 1024);
      case 3: return persistent_grid(k_sweep_x_f32_batch<2, 4, 4, B, A>, 256);
    }
  } else {
    switch (variant) {
      case 0: return persistent_grid(k_sweep_y_f32_batch<2, 2, B, A>, 128);
      case 1: return persistent_grid(k_sweep_y_f32_batch<2, 4, B, A>, 256);
      case 2: return persistent_grid(k_sweep_y_f32_batch<4, 2, B, A>, 128);
      case 3: return persistent_grid(k_sweep_y_f32_batch<4, 4, B, A>, 256);
    }
  }
  return 0;
}
}  // namespace

int f32_batch_grid(int axis, int variant, bool bary, bool adia) {
  if (bary) return adia ? batch_grid<true, true>(axis, variant) : batch_grid<true, false>(axis, variant);
  return adia ? batch_grid<false, true>(axis, variant) : batch_grid<false, false>(axis, variant);
}

hipError_t f32_sweep_x_batch(hipStream_t s, const DiskDev* dd, const DevF32* dfs, const int* order, int nslots,
                             int variant, int src, int grid, bool bary, bool adia) {
  switch (variant) {
    case 0: return sweep_x_batch_k<2, 4, 2>(s, dd, dfs, order, nslots, src, grid, bary, adia);
    case 1: return sweep_x_batch_k<2, 8, 1>(s, dd, dfs, order, nslots, src, grid, bary, adia);
    case 2: return sweep_x_batch_k<2, 16, 1>(s, dd, dfs, order, nslots, src, grid, bary, adia);
    case 3: return sweep_x_batch_k<2, 4, 4>(s, dd, dfs, order, nslots, src, grid, bary, adia);
  }
  return hipErrorInvalidValue;
}

hipError_t f32_sweep_y_batch(hipStream_t s, const DiskDev* dd, const DevF32* dfs, const int* base, int ndisk,
                             int variant, int src, int grid, bool bary, bool adia) {
  switch (variant) {
    case 0: return sweep_y_batch_k<2, 2>(s, dd, dfs, base, ndisk, src, grid, bary, adia);
    case 1: return sweep_y_batch_k<2, 4>(s, dd, dfs, base, ndisk, src, grid, bary, adia);
    case 2: return sweep_y_batch_k<4, 2>(s, dd, dfs, base, ndisk, src, grid, bary, adia);
    case 3: return sweep_y_batch_k<4, 4>(s, dd, dfs, base, ndisk, src, grid, bary, adia);
  }
  return hipErrorInvalidValue;
}

int f32_cfl_blocks(const DevGrid& g) { return red_blocks(g); }

hipError_t f32_cfl(const LaunchCtx& c, const DevF32& f, int src, int* n) {
  const int nb = red_blocks(c.g);
  if (nb > c.ncflp) return hipErrorInvalidValue;
  k_cfl_f32<<<nb, 256, 0, c.stream>>>(c.g, f, c.p, c.ss, c.cflp, src);
  *n = nb;
  return hipGetLastError();
}

hipError_t f32_reduce(const LaunchCtx& c, const DevF32& f, int src, int mode, double exclude, int slot) {
  const int nb = red_blocks(c.g);
  k_reduce_f32<<<nb, 256, 0, c.stream>>>(c.g, f, c.p, c.ss, src, mode, exclude, c.partials);
  return launch_sum_partials(c, nb, 1, &c.ss->red[slot]);
}

hipError_t f32_damp(const LaunchCtx& c, const DevF32& f, int src) {
  dim3 grid((c.g.jmax + 255) / 256, c.g.imax * c.g.kmax);
  k_damp_f32<<<grid, 256, 0, c.stream>>>(c.g, f, c.p, c.ss, src);
  return hipGetLastError();
}

hipError_t f32_fill_ghosts(const LaunchCtx& c, const DevF32& f, int src) {
  if (c.p.nlft == 3 && c.p.nrgh == 3) return hipSuccess;
  dim3 grid((c.g.jmax + 255) / 256, 2 * kNpad * c.g.kmax);
  k_fill_ghosts_f32<<<grid, 256, 0, c.stream>>>(c.g, f, c.p, src);
  return hipGetLastError();
}

hipError_t f32_totals(const LaunchCtx& c, const DevF32& f, int src, double* out2) {
  int nb = red_blocks(c.g);
  if (nb > kMaxPartials / 2) nb = kMaxPartials / 2;
  k_totals_f32<<<nb, 256, 0, c.stream>>>(c.g, f, src, c.partials);
  return launch_sum_partials(c, nb, 2, out2);
}

hipError_t f32_ring_means(const LaunchCtx& c, const DevF32& f, int src, int field, double* out) {
  if (field < 0 || field > 4) return hipErrorInvalidValue;
  k_ring_means_f32<<<c.g.imax * c.g.kmax, 256, 0, c.stream>>>(c.g, f, c.p.eos, src, field, out);
  return hipGetLastError();
}

}  // namespace pn
