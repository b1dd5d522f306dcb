// ppm_wave.hpp -- one directional PPMLR update of a 128-cell window on one wavefront.
//
// Method: the reference's ppmlr (ppmlr.cu:2-56) -- flatten, states (paraset, parabola,
// characteristic averages, interface force), Riemann solve, Lagrangian evolve, PPM remap --
// followed by the explicit viscosity (viscosity.cu:370-472).  Same operands in the same
// order per cell, so the FP64 instantiation reproduces the reference arithmetic.
//
// Execution model (not the reference's): one 64-lane wavefront owns a whole window, each
// lane K consecutive cells (cell c = K*lane + h).  Every stencil access is to the cell
// itself or to c+-1 / c+-2, i.e. to the same lane or a neighbouring lane, so neighbour
// values move by DPP wave shifts (v_mov_b32_dpp wave_shr:1 / wave_shl:1) instead of LDS
// round trips, and the ~50 block barriers of the reference window disappear.  All shifts
// run in wave-uniform control flow; per-cell index ranges of the reference stages become
// selects.
#pragma once

#include "pn_device.hpp"

namespace pn {

// ---------------------------------------------------------------------------------------
// lane exchange

// wave_shr:1 / wave_shl:1 with bound_ctrl: the lane without a source (0, resp. 63) reads 0,
// and the destination needs no initial value (one instruction per shift)
__device__ __forceinline__ int lane_prev_i(int x) {   // lane l receives lane l-1's value
  return __builtin_amdgcn_mov_dpp(x, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ int lane_next_i(int x) {   // lane l receives lane l+1's value
  return __builtin_amdgcn_mov_dpp(x, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ float lane_prev(float x) { return __int_as_float(lane_prev_i(__float_as_int(x))); }
__device__ __forceinline__ float lane_next(float x) { return __int_as_float(lane_next_i(__float_as_int(x))); }
__device__ __forceinline__ double lane_prev(double x) {
  int lo = lane_prev_i(__double2loint(x));
  int hi = lane_prev_i(__double2hiint(x));
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_next(double x) {
  int lo = lane_next_i(__double2loint(x));
  int hi = lane_next_i(__double2hiint(x));
  return __hiloint2double(hi, lo);
}

// per-lane slice of a window quantity
template <typename T, int K>
struct C {
  T a[K];
  __device__ __forceinline__ T& operator[](int h) { return a[h]; }
  __device__ __forceinline__ T operator[](int h) const { return a[h]; }
};

template <typename T, int K>
__device__ __forceinline__ C<T, K> splat(T s) {
  C<T, K> o;
#pragma unroll
  for (int h = 0; h < K; ++h) o[h] = s;
  return o;
}

// value of cell c-1
template <typename T, int K>
__device__ __forceinline__ C<T, K> prev1(const C<T, K>& x) {
  C<T, K> o;
  o[0] = lane_prev(x[K - 1]);
#pragma unroll
  for (int h = 1; h < K; ++h) o[h] = x[h - 1];
  return o;
}
// value of cell c+1
template <typename T, int K>
__device__ __forceinline__ C<T, K> next1(const C<T, K>& x) {
  C<T, K> o;
#pragma unroll
  for (int h = 0; h < K - 1; ++h) o[h] = x[h + 1];
  o[K - 1] = lane_next(x[0]);
  return o;
}
// value of cell c-2
template <typename T, int K>
__device__ __forceinline__ C<T, K> prev2(const C<T, K>& x) {
  if constexpr (K == 1) {
    return prev1(prev1(x));
  } else {
    C<T, K> o;
#pragma unroll
    for (int h = 0; h < K; ++h) o[h] = (h >= 2) ? x[h - 2] : lane_prev(x[K - 2 + h]);
    return o;
  }
}
// value of cell c+2
template <typename T, int K>
__device__ __forceinline__ C<T, K> next2(const C<T, K>& x) {
  if constexpr (K == 1) {
    return next1(next1(x));
  } else {
    C<T, K> o;
#pragma unroll
    for (int h = 0; h < K; ++h) o[h] = (h + 2 < K) ? x[h + 2] : lane_next(x[h + 2 - K]);
    return o;
  }
}

__device__ __forceinline__ int cell_of(int K, int h) { return K * (int)(threadIdx.x & 63) + h; }
// the cell index within a window when a wavefront holds 64 / SL windows side by side, SL
// lanes each (SL = 64: one window per wavefront, cell_of; lanes past the last whole window
// count from 0 again and store nothing).  Stencils reach across the seam between two windows
// exactly as they reach past a window's last cell: every stage's index range keeps such
// values out of the cells a window stores.
template <int SL>
__device__ __forceinline__ int cell_in(int K, int h) { return K * (int)((threadIdx.x & 63) % SL) + h; }

template <typename T>
__device__ __forceinline__ T min3(T A, T B, T Cc) {   // simple_func.cu:2-14
  if (A < B) return (Cc < A) ? Cc : A;
  return (Cc < B) ? Cc : B;
}

// the limited slope of parabola.cu:87-96: sign(t1) min(|t1|, 2 |da_l|, 2 |da_r|), zero at an
// extremum.  FP64 keeps the reference's compare chain (min3) operation for operation; FP32 takes
// the same minimum with one v_min3 and the sign by copysign (the values are the same: a
// minimum and a sign change round nothing).
template <typename T>
__device__ __forceinline__ T limited_slope(T t1, T da_l, T da_r) {
  if constexpr (sizeof(T) == 4) {
    t1 = copysignf(fminf(fminf(fabsf(t1), 2.0f * fabsf(da_l)), 2.0f * fabsf(da_r)), t1);
  } else {
    if (t1 < T(0.0)) t1 = -min3(-t1, T(2.0) * fabs(da_l), T(2.0) * fabs(da_r));
    else t1 = min3(t1, T(2.0) * fabs(da_l), T(2.0) * fabs(da_r));
  }
  if (da_l * da_r < T(0.0)) t1 = T(0.0);
  return t1;
}
template <typename T>
__device__ __forceinline__ T max3(T A, T B, T Cc) {   // simple_func.cu:16-28
  if (A > B) return (Cc > A) ? Cc : A;
  return (Cc > B) ? Cc : B;
}

// ---------------------------------------------------------------------------------------
// planet + star gravity (forces.cu:4-182).  Without bary: the star at the origin (its -1/r^2
// added by the caller in cylindrical r, included in the spherical forms) and the planet's
// indirect term; with bary (bary_flag): star at -x m/(1+m) and planet at x/(1+m) about their
// barycentre, no indirect term.

struct Planet {
  double m, x, y, vx, vy, rs;
  int bary;
};

template <typename T>
struct BaryPos {   // stm, plm and the two bodies' distances from the barycentre
  T stm, plm, plx, stx;
  __device__ __forceinline__ BaryPos(const Planet& p) {
    stm = T(1.0) / (T(1.0) + (T)p.m);
    plm = (T)p.m * stm;
    plx = (T)p.x * stm;
    stx = (T)p.x * plm;
  }
};

template <typename T>
__device__ __forceinline__ T grav_rad_cyl(const Planet& p, T rad, T azi, T z, T dt) {   // forces.cu:112-134
  T cosfac = cos(azi - ((T)p.y + dt * (T)p.vy));
  T rs = (T)p.rs;
  if (p.bary) {
    const BaryPos<T> b(p);
    T Rp = sqrt(rad * rad + b.plx * b.plx - T(2.0) * b.plx * rad * cosfac + z * z + rs * rs);
    T Rs = sqrt(rad * rad + b.stx * b.stx + T(2.0) * b.stx * rad * cosfac + z * z);
    return -b.plm * (rad - b.plx * cosfac) / (Rp * Rp * Rp) - b.stm * (rad + b.stx * cosfac) / (Rs * Rs * Rs);
  }
  T plm = (T)p.m, plx = (T)p.x;
  T Rp = sqrt(rad * rad + plx * plx - T(2.0) * plx * rad * cosfac + z * z + rs * rs);
  return -plm * (rad - plx * cosfac) / (Rp * Rp * Rp) - plm * cosfac / (plx * plx);
}
template <typename T>
__device__ __forceinline__ T grav_azi_cyl(const Planet& p, T rad, T azi, T z, T dt) {   // forces.cu:136-159
  T ph = azi - ((T)p.y + dt * (T)p.vy);
  T cosfac = cos(ph);
  T sinfac = sin(ph);
  T rs = (T)p.rs;
  if (p.bary) {
    const BaryPos<T> b(p);
    T Rp = sqrt(rad * rad + b.plx * b.plx - T(2.0) * b.plx * rad * cosfac + z * z + rs * rs);
    T Rs = sqrt(rad * rad + b.stx * b.stx + T(2.0) * b.stx * rad * cosfac + z * z);
    return -b.plm * b.plx * sinfac / (Rp * Rp * Rp) + b.stm * b.stx * sinfac / (Rs * Rs * Rs);
  }
  T plm = (T)p.m, plx = (T)p.x;
  T Rp = sqrt(rad * rad + plx * plx - T(2.0) * plx * rad * cosfac + z * z + rs * rs);
  return -plm * plx * sinfac / (Rp * Rp * Rp) + plm * sinfac / plx / plx;
}
template <typename T>
__device__ __forceinline__ T grav_pol_cyl(const Planet& p, T rad, T azi, T z, T dt) {   // forces.cu:161-182
  T cosfac = cos(azi - ((T)p.y + dt * (T)p.vy));
  T rs = (T)p.rs;
  if (p.bary) {
    const BaryPos<T> b(p);
    T Rp = sqrt(rad * rad + b.plx * b.plx - T(2.0) * b.plx * rad * cosfac + z * z + rs * rs);
    T Rs = sqrt(rad * rad + b.stx * b.stx + T(2.0) * b.stx * rad * cosfac + z * z);
    return -b.plm * z / (Rp * Rp * Rp) - b.stm * z / (Rs * Rs * Rs);
  }
  T plm = (T)p.m, plx = (T)p.x;
  T Rp = sqrt(rad * rad + plx * plx - T(2.0) * plx * rad * cosfac + z * z + rs * rs);
  return -plm * z / (Rp * Rp * Rp);
}
template <typename T>
__device__ __forceinline__ T grav_rad_sph(const Planet& p, T rad, T azi, T pol, T dt) {   // forces.cu:23-50
  T cosfac = cos(azi - ((T)p.y + dt * (T)p.vy));
  T sinpol = sin(pol);
  T rs = (T)p.rs;
  if (p.bary) {
    const BaryPos<T> b(p);
    T Rp = rad * rad + b.plx * b.plx - T(2.0) * b.plx * rad * cosfac * sinpol;
    Rp = sqrt(Rp + rs * rs);
    T Rs = sqrt(rad * rad + b.stx * b.stx + T(2.0) * b.stx * rad * cosfac * sinpol);
    return -b.plm * (rad - b.plx * cosfac * sinpol) / (Rp * Rp * Rp) -
           b.stm * (rad + b.stx * cosfac * sinpol) / (Rs * Rs * Rs);
  }
  T stm = T(1.0) / (T(1.0) + (T)p.m);
  T plm = (T)p.m * stm;
  T plx = (T)p.x;
  T Rp = rad * rad + plx * plx - T(2.0) * plx * rad * cosfac * sinpol;
  Rp = sqrt(Rp + rs * rs);
  return -plm * (rad - plx * cosfac * sinpol) / (Rp * Rp * Rp) - stm / (rad * rad) -
         plm * sinpol * cosfac / (plx * plx);
}
template <typename T>
__device__ __forceinline__ T grav_azi_sph(const Planet& p, T rad, T azi, T pol, T dt) {   // forces.cu:52-81
  T ph = azi - ((T)p.y + dt * (T)p.vy);
  T cosfac = cos(ph);
  T sinfac = sin(ph);
  T sinpol = sin(pol);
  rad /= sinpol;
  T rs = (T)p.rs;
  if (p.bary) {
    const BaryPos<T> b(p);
    T Rp = rad * rad + b.plx * b.plx - T(2.0) * b.plx * rad * cosfac * sinpol;
    Rp = sqrt(Rp + rs * rs);
    T Rs = sqrt(rad * rad + b.stx * b.stx + T(2.0) * b.stx * rad * cosfac * sinpol);
    return -b.plm * b.plx * sinfac / (Rp * Rp * Rp) + b.stm * b.stx * sinfac / (Rs * Rs * Rs);
  }
  T stm = T(1.0) / (T(1.0) + (T)p.m);
  T plm = (T)p.m * stm;
  T plx = (T)p.x;
  T Rp = rad * rad + plx * plx - T(2.0) * plx * rad * cosfac * sinpol;
  Rp = sqrt(Rp + rs * rs);
  return -plm * plx * sinfac / (Rp * Rp * Rp) + plm * sinpol * sinfac / plx / plx;
}
template <typename T>
__device__ __forceinline__ T grav_pol_sph(const Planet& p, T rad, T azi, T pol, T dt) {   // forces.cu:83-110
  T cosfac = cos(azi - ((T)p.y + dt * (T)p.vy));
  T cospol = cos(pol);
  T sinpol = sin(pol);
  T rs = (T)p.rs;
  if (p.bary) {
    const BaryPos<T> b(p);
    T Rp = rad * rad + b.plx * b.plx - T(2.0) * b.plx * rad * cosfac * sinpol;
    Rp = sqrt(Rp + rs * rs);
    T Rs = sqrt(rad * rad + b.stx * b.stx + T(2.0) * b.stx * rad * cosfac * sinpol);
    return b.plm * b.plx * cosfac * cospol / (Rp * Rp * Rp) - b.stm * b.stx * cosfac * cospol / (Rs * Rs * Rs);
  }
  T stm = T(1.0) / (T(1.0) + (T)p.m);
  T plm = (T)p.m * stm;
  T plx = (T)p.x;
  T Rp = rad * rad + plx * plx - T(2.0) * plx * rad * cosfac * sinpol;
  Rp = sqrt(Rp + rs * rs);
  return plm * plx * cosfac * cospol / (Rp * Rp * Rp) - plm * cosfac * cospol / (sinpol * plx * plx);
}
template <typename T>
__device__ __forceinline__ T potential(bool planet_on, const Planet& p, T rad, T azi, T z) {   // forces.cu:4-21
  if (planet_on) {
    T cosfac = cos(azi - (T)p.y);
    T stm = T(1.0) / (T(1.0) + (T)p.m);
    T plm = (T)p.m * stm;
    T plx = (T)p.x * stm;
    T stx = (T)p.x * plm;
    T rs = (T)p.rs;
    T Rp = rad * rad + plx * plx - T(2.0) * plx * rad * cosfac + z * z;
    Rp = sqrt(Rp + rs * rs);
    T Rs = sqrt(rad * rad + stx * stx + T(2.0) * stx * rad * cosfac + z * z);
    return -plm / Rp - stm / Rs;
  }
  return T(-1.0) / sqrt(rad * rad + z * z);
}

// ---------------------------------------------------------------------------------------
// window configuration

template <int AXIS_, int GX_, int GY_, int GZ_, int EOS_, int VISC_>
struct Cfg {
  static constexpr int AXIS = AXIS_, GX = GX_, GY = GY_, GZ = GZ_, EOS = EOS_, VISC = VISC_;
};

// the energy a cell carries into the Lagrangian step from the window's p, rho, u and first
// transverse slot v: EOS 2 as ppm.cu:106-108 has it (v already scaled by r_cyl in the radial
// sweep), EOS 3 the internal energy plus the sweep direction's kinetic energy; 0 otherwise
template <class CF>
__device__ __forceinline__ double energy_of(double p, double r, double u, double v, double gam) {
  if constexpr (CF::EOS == 2) return p / (r * (gam - 1.0)) + 0.5 * ((u * u) + (v * v));
  else if constexpr (CF::EOS == 3) return p / (r * (gam - 1.0)) + 0.5 * (u * u);
  else return 0.0;
}

struct RunFlags {
  int flat;      // flattening on
  int planet;    // plnt_flag > 0
  int remap_v;   // the sweep's first transverse slot is carried (else it is identically 0)
  int remap_w;   // the sweep's second transverse slot is carried
  int ndim;
};

template <typename T, int K>
struct Para {
  C<T, K> c0, c1, c2, c3, c4;
};

// paraset (parabola.cu:138-176) on the index range [nmin, nmaxp)
template <typename T, int K, int SL = 64>
__device__ __forceinline__ Para<T, K> paraset(int nmin, int nmaxp, const C<T, K>& dx) {
  Para<T, K> P;
  C<T, K> dxp = next1(dx);
  C<T, K> an, ain, bn, bi, cc, ci;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    an[h] = dx[h] + dxp[h];
    ain[h] = T(1.0) / an[h];
    bn[h] = an[h] + dx[h];
    bi[h] = T(1.0) / bn[h];
    cc[h] = an[h] + dxp[h];
    ci[h] = T(1.0) / cc[h];
  }
  C<T, K> anm = prev1(an), anp = next1(an), bnm = prev1(bn), ainm = prev1(ain);
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const int c = cell_in<SL>(K, h);
    T c0 = T(0), c1 = T(0), c2 = T(0), c3 = T(0), c4 = T(0);
    if (c >= nmin + 1 && c < nmaxp - 2) {
      T d = T(1.0) / (anm[h] + anp[h]);
      c0 = dx[h] * ain[h] * (T(1.0) + T(2.0) * d * dxp[h] * (anm[h] * bi[h] - anp[h] * ci[h]));
      c1 = -d * dx[h] * anm[h] * bi[h];
      c2 = d * dxp[h] * anp[h] * ci[h];
    }
    if (c >= nmin + 1 && c < nmaxp - 1) {
      T d = dx[h] / (anm[h] + dxp[h]);
      c3 = d * bnm[h] * ain[h];
      c4 = d * cc[h] * ainm[h];
    }
    P.c0[h] = c0; P.c1[h] = c1; P.c2[h] = c2; P.c3[h] = c3; P.c4[h] = c4;
  }
  return P;
}

template <typename T, int K>
struct Parab {
  C<T, K> al, da, a6;   // left value, jump, curvature; valid on [nmin+2, nmaxp-2)
};

// parabola (parabola.cu:70-136) on [nmin, nmaxp)
template <typename T, int K, int SL = 64>
__device__ __forceinline__ Parab<T, K> parabola(int nmin, int nmaxp, const C<T, K>& flat, const C<T, K>& val,
                                                const Para<T, K>& P) {
  C<T, K> vm = prev1(val), vp = next1(val);
  C<T, K> slope, dar;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    T a = val[h];
    T da_r = vp[h] - a;
    T da_l = a - vm[h];
    slope[h] = limited_slope<T>(P.c3[h] * da_r + P.c4[h] * da_l, da_l, da_r);
    dar[h] = da_r;
  }
  C<T, K> slp = next1(slope);
  C<T, K> ar;
#pragma unroll
  for (int h = 0; h < K; ++h) ar[h] = val[h] + P.c0[h] * dar[h] + P.c1[h] * slp[h] + P.c2[h] * slope[h];
  C<T, K> alv = prev1(ar);
  Parab<T, K> o;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const int c = cell_in<SL>(K, h);
    T a = val[h];
    T t1 = flat[h] * a;
    T t2 = T(1.0) - flat[h];
    T arr = t1 + t2 * ar[h];
    T la = t1 + t2 * alv[h];
    t1 = arr - la;
    t2 = t1 * t1;
    t1 = t1 * (T(6.0) * (a - T(0.5) * (arr + la)));
    if ((arr - a) * (a - la) <= T(0.0)) {
      arr = a;
      la = a;
    }
    if (t2 < t1) la = T(3.0) * a - T(2.0) * arr;
    if (t2 < -t1) arr = T(3.0) * a - T(2.0) * la;
    bool in = (c >= nmin + 2 && c < nmaxp - 2);
    o.da[h] = in ? arr - la : T(0);
    o.a6[h] = in ? T(6.0) * (a - T(0.5) * (arr + la)) : T(0);
    o.al[h] = in ? la : T(0);
  }
  return o;
}

// derivs (viscosity.cu:2-19): first and second derivative of a quadratic through the
// left-edge coordinates of cells c-1, c, c+1
template <typename T>
__device__ __forceinline__ void derivs(T xm, T x0, T xp, T um, T u0, T up, T& first, T& second) {
  T x2 = x0 - xm;
  T x3 = xp - xm;
  T h2 = (u0 - um) / x2;
  T h3 = (up - um) / x3;
  T A = (h3 - h2) / (x3 - x2);
  T B = (x3 * h2 - x2 * h3) / (x3 - x2);
  second = T(2.0) * A;
  first = second * x2 + B;
}

// Window state and per-cell geometry.
template <typename T, int K>
struct Win {
  C<T, K> r, p, u, v, w, e;
  C<T, K> xa0, dx0, dvol0, rad, azi, pol, vrot;
};

// interface / centre force of one axis (forces.cu:185-372)
template <typename T, class CF>
__device__ __forceinline__ T force_x(const RunFlags& rf, const Planet& pl, bool bound, T dt, T rad, T azi, T pol,
                                     T vc, T vm, T wc, T wm) {
  T grav, ftot;
  if constexpr (CF::GX == 1) {
    if (rf.planet) {
      grav = grav_rad_cyl<T>(pl, rad, azi, pol, dt);
      if (!pl.bary) grav += T(-1.0) / rad / rad;
    } else {
      T dis = sqrt(rad * rad + pol * pol);
      grav = -rad / dis / dis / dis;
    }
    ftot = grav;
    if (bound) {
      T l = T(0.5) * (vc + vm);
      ftot += l * l / rad / rad / rad;
    } else {
      ftot += vc * vc / rad / rad / rad;
    }
  } else if constexpr (CF::GX == 2) {
    grav = rf.planet ? grav_rad_sph<T>(pl, rad, azi, pol, dt) : T(-1.0) / rad / rad;
    ftot = grav;
    T sp = rad * sin(pol);
    if (bound) {
      T lw = T(0.5) * (wc + wm);
      T lv = T(0.5) * (vc + vm);
      ftot += lw * lw / rad / rad / rad;
      ftot += lv * lv / rad / (sp * sp);
    } else {
      ftot += wc * wc / rad / rad / rad;
      ftot += vc * vc / rad / (sp * sp);
    }
  } else {
    ftot = T(0.0);
  }
  return ftot;
}

template <typename T, class CF>
__device__ __forceinline__ T force_y(const RunFlags& rf, const Planet& pl, T dt, T rad, T azi, T pol) {
  if constexpr (CF::GY == 3) return rf.planet ? grav_azi_cyl<T>(pl, rad, azi, pol, dt) : T(0.0);
  else if constexpr (CF::GY == 4) return rf.planet ? grav_azi_sph<T>(pl, rad, azi, pol, dt) : T(0.0);
  else return T(0.0);
}

template <typename T, class CF>
__device__ __forceinline__ T force_z(const RunFlags& rf, const Planet& pl, bool bound, T dt, T rad, T azi, T pol,
                                     T wc, T wm) {
  if constexpr (CF::GZ == 0) {
    if (rf.planet) return grav_pol_cyl<T>(pl, rad, azi, pol, dt);
    T dis = sqrt(rad * rad + pol * pol);
    return -pol / dis / dis / dis;
  } else if constexpr (CF::GZ == 5) {
    T grav = rf.planet ? grav_pol_sph<T>(pl, rad, azi, pol, dt) : T(0.0);
    T rc = rad * sin(pol);
    T l = bound ? T(0.5) * (wc + wm) : wc;
    return grav + (l * l * cos(pol)) / rc / rc / rc;
  } else {
    return T(0.0);
  }
}

// One PPMLR update of the window (ppmlr.cu:2-56).  nmax = window length.  On return the
// cells [6, nmax-6) hold the updated r, u, v, w (and p for EOS 1, 2).
template <typename T, int K, class CF>
__device__ __forceinline__ void ppmlr_wave(Win<T, K>& W, int nmax, T dt, T FrRot, const Planet& pl,
                                           const RunFlags& rf, T gam) {
  constexpr int AX = CF::AXIS;
  const T hdt = T(0.5) * dt;
  const T smallp = T(1.0e-15), smallr = T(1.0e-15);
  const T fourthd = T(1.3333333333333333333333333);
  const T third = T(0.3333333333333333333333333);
  const T gamm = gam - T(1.0);

  // ---- flatten (flatten.cu:7-40)
  C<T, K> flat = splat<T, K>(T(0));
  if (rf.flat) {
    C<T, K> pm1 = prev1(W.p), pp1 = next1(W.p), pm2 = prev2(W.p), pp2 = next2(W.p);
    C<T, K> um1 = prev1(W.u), up1 = next1(W.u);
    C<T, K> steep;
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const int c = cell_of(K, h);
      T s = T(0.0);
      if (c >= 2 && c < nmax - 2) {
        T dp1 = pp1[h] - pm1[h];
        T dp2 = pp2[h] - pm2[h];
        if (fabs(dp2) < smallp) dp2 = smallp;
        T ddp = fabs(dp1) / fmin(pp1[h], pm1[h]);
        bool shock = ddp > T(0.33);
        if (um1[h] < up1[h]) shock = false;
        if (shock) s = fmax(T(0.0), (dp1 / dp2 - T(0.75)) * T(5.0));
      }
      steep[h] = s;
    }
    C<T, K> sm = prev1(steep), sp = next1(steep);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const int c = cell_of(K, h);
      flat[h] = (c >= 2 && c < nmax - 2) ? fmin(T(0.5), max3(sm[h], steep[h], sp[h])) : T(0.0);
    }
  }

  // ---- states (states.cu:3-73)
  Para<T, K> P0 = paraset(0, nmax, W.dx0);
  C<T, K> ftot_s;
  {
    C<T, K> vm = prev1(W.v);
    C<T, K> wm;
    if constexpr (CF::GX == 2 || CF::GZ == 5) wm = prev1(W.w);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      T f = T(0);
      if constexpr (AX == 0) {
        T wmh = (CF::GX == 2) ? wm[h] : T(0);
        f = force_x<T, CF>(rf, pl, true, hdt, W.xa0[h], W.azi[h] + hdt * W.vrot[h], W.pol[h], W.v[h], vm[h], W.w[h],
                           wmh);
      } else if constexpr (AX == 1) {
        f = force_y<T, CF>(rf, pl, hdt, W.rad[h], W.xa0[h] + hdt * W.vrot[h], W.pol[h]);
      } else {
        T wmh = (CF::GZ == 5) ? wm[h] : T(0);
        f = force_z<T, CF>(rf, pl, true, hdt, W.rad[h], W.azi[h] + hdt * W.vrot[h], W.xa0[h], W.w[h], wmh);
      }
      ftot_s[h] = f;
    }
  }
  C<T, K> cs, csfac;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    T s = hdt * sqrt(gam * W.p[h] / W.r[h]) / W.dx0[h];
    if constexpr (AX == 1) s /= W.rad[h];
    if constexpr (AX == 2 && CF::GZ == 5) s /= W.rad[h];
    cs[h] = s;
    csfac[h] = T(1.0) - fourthd * s;
  }
  // left state at interface c = right-edge average of cell c-1; right state = left-edge
  // average of cell c (states.cu:45-68)
  C<T, K> Spl, Spr, Srl, Srr, Sul, Sur;
  {
    Parab<T, K> q = parabola(0, nmax, flat, W.p, P0);
    C<T, K> re, le;
#pragma unroll
    for (int h = 0; h < K; ++h) {
      re[h] = q.al[h] + q.da[h] - cs[h] * (q.da[h] - csfac[h] * q.a6[h]);
      le[h] = q.al[h] + cs[h] * (q.da[h] + csfac[h] * q.a6[h]);
    }
    C<T, K> rem = prev1(re);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      Spl[h] = fmax(smallp, rem[h]);
      Spr[h] = fmax(smallp, le[h]);
    }
  }
  {
    Parab<T, K> q = parabola(0, nmax, flat, W.r, P0);
    C<T, K> re, le;
#pragma unroll
    for (int h = 0; h < K; ++h) {
      re[h] = q.al[h] + q.da[h] - cs[h] * (q.da[h] - csfac[h] * q.a6[h]);
      le[h] = q.al[h] + cs[h] * (q.da[h] + csfac[h] * q.a6[h]);
    }
    C<T, K> rem = prev1(re);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      Srl[h] = fmax(smallr, rem[h]);
      Srr[h] = fmax(smallr, le[h]);
    }
  }
  {
    Parab<T, K> q = parabola(0, nmax, flat, W.u, P0);
    C<T, K> re, le;
#pragma unroll
    for (int h = 0; h < K; ++h) {
      re[h] = q.al[h] + q.da[h] - cs[h] * (q.da[h] - csfac[h] * q.a6[h]);
      le[h] = q.al[h] + cs[h] * (q.da[h] + csfac[h] * q.a6[h]);
    }
    C<T, K> rem = prev1(re);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      Sul[h] = rem[h] + hdt * ftot_s[h];
      Sur[h] = le[h] + hdt * ftot_s[h];
    }
  }

  // ---- riemann (riemann.cu:3-115)
  C<T, K> umid, pmid;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    T pl_ = Spl[h], rl = Srl[h], ul = Sul[h], pr = Spr[h], rr = Srr[h], ur = Sur[h];
    T um, pm;
    if constexpr (CF::EOS == 0) {
      T t1 = T(0.5) * (sqrt(pl_ / rl) + sqrt(pr / rr));
      T t2 = T(-0.5) * (ur - ul) / t1;
      pm = sqrt(pl_ * pr) * exp(t2);
      um = ul - t1 * log(pm / pl_);
    } else if constexpr (CF::EOS == 2 || CF::EOS == 3) {
      const T gamfac2 = gam + T(1.0), gamz = gamm / gam / T(2.0);
      T cs_l = sqrt(gam * pl_ / rl);
      T cs_r = sqrt(gam * pr / rr);
      pm = pow((cs_l + cs_r - (gamm / T(2.0)) * (ur - ul)) / (cs_l * pow(pl_, -gamz) + cs_r * pow(pr, -gamz)),
               T(1.0) / gamz);
      T t1, t2;
      if (pm > pr) t1 = (pm - pr) * sqrt(T(2.0) / (rr * (gamfac2 * pm + pr * gamm)));
      else t1 = T(2.0) * cs_r * (pow(pm / pr, gamz) - T(1.0)) / gamm;
      if (pm > pl_) t2 = (pm - pl_) * sqrt(T(2.0) / (rl * (gamfac2 * pm + pl_ * gamm)));
      else t2 = T(2.0) * cs_l * (pow(pm / pl_, gamz) - T(1.0)) / gamm;
      um = T(0.5) * (ur + ul) + T(0.5) * (t1 - t2);
    } else {
      const T gamfac2 = gam + T(1.0), gamfac1 = gamfac2 / gam / T(2.0);
      T cs_l = sqrt(gam * pl_ * rl);
      T cs_r = sqrt(gam * pr * rr);
      rl = T(1.0) / rl;
      rr = T(1.0) / rr;
      pm = pr - pl_ - cs_r * (ur - ul);
      pm = pl_ + pm * cs_l / (cs_l + cs_r);
      pm = fmax(smallp, pm);
      T t1 = T(0), t2 = T(0);
      for (int it = 0; it < 10; ++it) {
        T dum = pm;
        t1 = T(1.0) + gamfac1 * (pm - pl_) / pl_;
        t2 = T(1.0) + gamfac1 * (pm - pr) / pr;
        t1 = cs_l * sqrt(t1);
        t2 = cs_r * sqrt(t2);
        T en_l = T(4.0) * rl * t1 * t1;
        T en_r = T(4.0) * rr * t2 * t2;
        en_l = -en_l * t1 / (en_l - gamfac2 * (pm - pl_));
        en_r = en_r * t2 / (en_r - gamfac2 * (pm - pr));
        T uml = ul - (pm - pl_) / t1;
        T umr = ur + (pm - pr) / t2;
        pm = pm + (umr - uml) * (en_r * en_l) / (en_r - en_l);
        pm = fmax(smallp, pm);
        if (fabs(pm - dum) / pm < T(1.0e-8)) break;
      }
      T uml = ul - (pm - pl_) / t1;
      T umr = ur + (pm - pr) / t2;
      um = T(0.5) * (uml + umr);
    }
    umid[h] = um;
    pmid[h] = fmax(smallp, pm);
  }

  // ---- evolve (evolve.cu:3-166)
  C<T, K> dtdm, vf, xa, dxl, dvol;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const int c = cell_of(K, h);
    T delx;
    if constexpr (AX == 0) {
      delx = dt * umid[h];
    } else if constexpr (AX == 1) {
      if constexpr (CF::GY > 2) delx = dt * (umid[h] / W.rad[h] - W.vrot[h] - FrRot);
      else delx = dt * umid[h];
    } else {
      if constexpr (CF::GZ == 5) delx = dt * umid[h] / W.rad[h];
      else delx = dt * umid[h];
    }
    dtdm[h] = dt / (W.r[h] * W.dvol0[h]);
    bool in = (c >= 3 && c < nmax - 2);
    vf[h] = in ? W.xa0[h] + delx : W.xa0[h] + T(0.5) * W.dx0[h];
  }
  {
    C<T, K> vfp = next1(vf);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const int c = cell_of(K, h);
      bool in = (c >= 3 && c < nmax - 3);
      dxl[h] = in ? vfp[h] - vf[h] : W.dx0[h];
      xa[h] = in ? vf[h] : W.xa0[h];
    }
  }
  C<T, K> fto1, fto2, xcl;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    T xc_ = xa[h] + T(0.5) * dxl[h];
    xcl[h] = xc_;
    if constexpr (AX == 0) {
      fto1[h] = force_x<T, CF>(rf, pl, false, T(0.0), W.rad[h], W.azi[h], W.pol[h], W.v[h], T(0), W.w[h], T(0));
      fto2[h] = force_x<T, CF>(rf, pl, false, dt, xc_, W.azi[h] + dt * W.vrot[h], W.pol[h], W.v[h], T(0), W.w[h], T(0));
    } else if constexpr (AX == 1) {
      fto1[h] = force_y<T, CF>(rf, pl, T(0.0), W.rad[h], W.azi[h], W.pol[h]);
      fto2[h] = force_y<T, CF>(rf, pl, dt, W.rad[h], xc_ + dt * W.vrot[h], W.pol[h]);
    } else {
      fto1[h] = force_z<T, CF>(rf, pl, false, T(0.0), W.rad[h], W.azi[h], W.pol[h], W.w[h], T(0));
      fto2[h] = force_z<T, CF>(rf, pl, false, dt, W.rad[h], W.azi[h] + dt * W.vrot[h], xc_, W.w[h], T(0));
    }
  }
  C<T, K> umidp;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const T x_ = xa[h], x0 = W.xa0[h], d_ = dxl[h];
    umidp[h] = umid[h] * pmid[h];
    T dv, vv;
    if constexpr (AX == 0) {
      if constexpr (CF::GX == 0) { dv = d_; vv = T(1.0); }
      else if constexpr (CF::GX == 1) { dv = d_ * (x_ + T(0.5) * d_); vv = T(0.5) * (x_ + x0); }
      else { dv = d_ * (x_ * (x_ + d_) + d_ * d_ * third); vv = (x_ - x0) * (third * (x_ - x0) + x0) + (x0 * x0); }
    } else if constexpr (AX == 1) {
      if constexpr (CF::GY == 0) { dv = d_; vv = T(1.0); }
      else { dv = d_ * W.rad[h]; vv = T(1.0); }
    } else {
      if constexpr (CF::GZ == 0) { dv = d_; vv = T(1.0); }
      else {
        T delx = x_ - x0;
        T dum = cos(x_);
        dv = (dum - cos(x_ + d_)) * W.rad[h];
        vv = (delx == T(0.0)) ? sin(x_) : (cos(x0) - dum) / delx;
      }
    }
    dvol[h] = dv;
    vf[h] = vv;   // area factor at the left face
  }
  C<T, K> q;
  {
    C<T, K> pmp = next1(pmid), vfp = next1(vf), ump = next1(umidp);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const int c = cell_of(K, h);
      bool in = (c >= 3 && c < nmax - 3);
      T qh = T(0);
      if constexpr (CF::EOS == 0) qh = W.p[h] / W.r[h];
      else if constexpr (CF::EOS == 1) qh = W.p[h] / pow(W.r[h], gam);
      T rn = fmax(W.r[h] * (W.dvol0[h] / dvol[h]), smallr);
      T delx = T(0.5) * dt * (fto1[h] + fto2[h]);
      delx += -dtdm[h] * (pmp[h] - pmid[h]) * T(0.5) * (vfp[h] + vf[h]);
      T un = W.u[h] + delx;
      if constexpr (CF::EOS == 2) {
        T ef = W.e[h] - dtdm[h] * (vfp[h] * ump[h] - vf[h] * umidp[h]);
        T x_ = xcl[h];
        if constexpr (AX == 0)
          qh = ef - T(0.5) * ((un * un) + (W.v[h] * W.v[h] / x_ / x_)) - potential<T>(rf.planet, pl, x_, W.azi[h], W.pol[h]);
        else if constexpr (AX == 1)
          qh = ef - T(0.5) * (un * un) - potential<T>(rf.planet, pl, W.rad[h], x_, W.pol[h]);
        else
          qh = ef - T(0.5) * (un * un) - potential<T>(rf.planet, pl, W.rad[h], W.azi[h], x_);
        qh = fmax(gamm * qh, smallp);
      } else if constexpr (CF::EOS == 3) {
        // the energy-consistent form (the commented line of evolve.cu:146): e = p / (gamm rho)
        // + u^2 / 2 gains the pressure work and the time-centred work of the forces that move
        // u; the internal energy is what remains after the new u^2 / 2
        T ef = W.e[h] - dtdm[h] * (vfp[h] * ump[h] - vf[h] * umidp[h]) +
               T(0.5) * dt * (W.u[h] * fto1[h] + un * fto2[h]);
        qh = fmax(gamm * (ef - T(0.5) * (un * un)), smallp);
      }
      q[h] = qh;
      if (in) {
        W.r[h] = rn;
        W.u[h] = un;
      }
    }
  }

  // ---- remap (remap.cu:14-226)
  Para<T, K> P1 = paraset(3, nmax - 3, dxl);
  C<T, K> dxf, dxf2, fr, delC;
  C<T, K> dxm = prev1(dxl);
  C<T, K> lvol;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const T x_ = xa[h], x0 = W.xa0[h];
    T lv;
    if constexpr (AX == 0) {
      if constexpr (CF::GX == 0) lv = x_ - x0;
      else if constexpr (CF::GX == 1) { lv = x_ - x0; lv *= (x0 + T(0.5) * lv); }
      else { lv = x_ - x0; lv *= x0 * x_ + third * (lv * lv); }
    } else if constexpr (AX == 1) {
      if constexpr (CF::GY == 0) lv = x_ - x0;
      else lv = (x_ - x0) * W.rad[h];
    } else {
      if constexpr (CF::GZ == 0) lv = x_ - x0;
      else lv = (cos(x0) - cos(x_)) * W.rad[h];
    }
    lvol[h] = lv;
    T d = x_ - x0;
    delC[h] = (d >= T(0.0)) ? T(1) : T(0);
    if (d >= T(0.0)) {
      d = T(0.5) * d / dxm[h];
      dxf2[h] = T(1.0) - fourthd * d;
    } else {
      d = T(0.5) * d / dxl[h];
      dxf2[h] = T(1.0) + fourthd * d;
    }
    dxf[h] = d;
  }
  // face average of a parabola swept through the face (remap.cu:59-90, get_flux :1-10)
  auto face = [&](const Parab<T, K>& Q, C<T, K>& out) {
    C<T, K> alm = prev1(Q.al), dam = prev1(Q.da), a6m = prev1(Q.a6);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      if (delC[h] != T(0)) out[h] = alm[h] + dam[h] - dxf[h] * (dam[h] - dxf2[h] * a6m[h]);
      else out[h] = Q.al[h] - dxf[h] * (Q.da[h] + dxf2[h] * Q.a6[h]);
    }
  };
  C<T, K> mass, invm;
  {
    Parab<T, K> Q = parabola(3, nmax - 3, flat, W.r, P1);
    C<T, K> fv;
    face(Q, fv);
#pragma unroll
    for (int h = 0; h < K; ++h) fr[h] = fv[h] * lvol[h];
    C<T, K> frp = next1(fr);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const int c = cell_of(K, h);
      T m = W.r[h] * dvol[h];
      T rn = fmax(smallr, (m + fr[h] - frp[h]) / W.dvol0[h]);
      mass[h] = m;
      invm[h] = T(1.0) / (rn * W.dvol0[h]);
      if (c >= 6 && c < nmax - 6) W.r[h] = rn;
    }
  }
  auto remap_field = [&](C<T, K>& X) {
    Parab<T, K> Q = parabola(3, nmax - 3, flat, X, P1);
    C<T, K> fl;
    face(Q, fl);
#pragma unroll
    for (int h = 0; h < K; ++h) fl[h] = fl[h] * fr[h];
    C<T, K> flp = next1(fl);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const int c = cell_of(K, h);
      if (c >= 6 && c < nmax - 6) X[h] = (X[h] * mass[h] + (fl[h] - flp[h])) * invm[h];
    }
  };
  remap_field(W.u);
  if (rf.remap_v) remap_field(W.v);
  if (rf.remap_w) remap_field(W.w);
  if constexpr (CF::EOS != 0) {
    remap_field(q);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const int c = cell_of(K, h);
      if (c >= 6 && c < nmax - 6) {
        if constexpr (CF::EOS == 1) W.p[h] = fmax(smallp, pow(W.r[h], gam) * q[h]);
        else W.p[h] = fmax(smallp, W.r[h] * q[h]);
      }
    }
  }
}

// device_viscosity_r (viscosity.cu:370-421); R = cylindrical radius per cell, x = left edges
template <typename T, int K>
__device__ __forceinline__ void visc_r(Win<T, K>& W, int lim, T dt, const C<T, K>& R, const C<T, K>& nu, bool do_w) {
  C<T, K> nr, vv;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    nr[h] = nu[h] * W.r[h];
    vv[h] = W.v[h] / (R[h] * R[h]);
  }
  C<T, K> xm = prev1(W.xa0), xp = next1(W.xa0);
  C<T, K> nm = prev1(nr), np = next1(nr);
  C<T, K> um = prev1(W.u), up = next1(W.u);
  C<T, K> vm = prev1(vv), vp = next1(vv);
  C<T, K> wm, wp;
  if (do_w) {
    wm = prev1(W.w);
    wp = next1(W.w);
  }
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const int c = cell_of(K, h);
    T rho = W.r[h], n = nr[h], Ri = R[h];
    T dn, d1, d2;
    derivs<T>(xm[h], W.xa0[h], xp[h], nm[h], n, np[h], dn, d2);
    derivs<T>(xm[h], W.xa0[h], xp[h], um[h], W.u[h], up[h], d1, d2);
    T ddu = T(2.0) * n * d2 + T(2.0) * dn * d1 + T(2.0) * n * d1 / Ri - T(2.0) * n * W.u[h] / Ri / Ri;
    ddu = W.u[h] + dt * ddu / rho;
    derivs<T>(xm[h], W.xa0[h], xp[h], vm[h], vv[h], vp[h], d1, d2);
    T ddv = Ri * n * d2 + Ri * dn * d1 + T(3.0) * n * d1;
    ddv = vv[h] * Ri + dt * ddv / rho;
    T ddw = W.w[h];
    if (do_w) {
      derivs<T>(xm[h], W.xa0[h], xp[h], wm[h], W.w[h], wp[h], d1, d2);
      T t = n * d2 + dn * d1 + n * d1 / Ri;
      ddw = W.w[h] + dt * t / rho;
    }
    if (c >= kNpad && c < lim - kNpad) {
      W.u[h] = ddu;
      W.v[h] = ddv * Ri;
      if (do_w) W.w[h] = ddw;
    }
  }
}

// device_viscosity_p (viscosity.cu:423-472), called with (u, v, w) = (u_r, v_phi, w) of the
// azimuthal sweep, i.e. sweep fields (w, u, v); x = left edges times R
template <typename T, int K>
__device__ __forceinline__ void visc_p(Win<T, K>& W, int lim, T dt, T R, T nu, const C<T, K>& x, bool do_w) {
  C<T, K> nr;
#pragma unroll
  for (int h = 0; h < K; ++h) nr[h] = nu * W.r[h];
  C<T, K> xm = prev1(x), xp = next1(x);
  C<T, K> nm = prev1(nr), np = next1(nr);
  // parameter v of the reference = sweep u (v_phi); parameter u = sweep w (u_r);
  // parameter w = sweep v
  C<T, K> pvm = prev1(W.u), pvp = next1(W.u);
  C<T, K> pum = prev1(W.w), pup = next1(W.w);
  C<T, K> pwm, pwp;
  if (do_w) {
    pwm = prev1(W.v);
    pwp = next1(W.v);
  }
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const int c = cell_of(K, h);
    T rho = W.r[h], n = nr[h];
    T dn, d1, d2;
    derivs<T>(xm[h], x[h], xp[h], nm[h], n, np[h], dn, d2);
    derivs<T>(xm[h], x[h], xp[h], pvm[h], W.u[h], pvp[h], d1, d2);
    T ddv = T(2.0) * n * d2 + T(2.0) * dn * d1;
    T ddu = T(0.0);
    ddv = W.u[h] + dt * ddv / rho;
    T ddw = W.v[h];
    if (do_w) {
      derivs<T>(xm[h], x[h], xp[h], pwm[h], W.v[h], pwp[h], d1, d2);
      T t = n * d2 + dn * d1;
      ddw = W.v[h] + dt * t / rho;
    }
    derivs<T>(xm[h], x[h], xp[h], pum[h], W.w[h], pup[h], d1, d2);
    ddu += n * d2 + dn * d1;
    ddu = W.w[h] + dt * ddu / rho;
    if (c >= kNpad && c < lim - kNpad) {
      W.u[h] = ddv;
      if (do_w) W.v[h] = ddw;
      W.w[h] = ddu;
    }
  }
}

// device_viscosity_z (viscosity.cu:474-518): nu rho with the constant vis_nu, the component
// along the pencil (the reference's w) taking 2 nu rho d2 + 2 dn d1 and the transverse ones
// nu rho d2 + dn d1, no curvature terms.  The theta sweep passes (u, v, w) = (v_r, v_phi,
// v_theta) and x = r times the theta left edges.
template <typename T, int K>
__device__ __forceinline__ void visc_z(const C<T, K>& rho, int lim, T dt, T nu, const C<T, K>& x, C<T, K>& u,
                                       C<T, K>& v, C<T, K>& w) {
  C<T, K> nr;
#pragma unroll
  for (int h = 0; h < K; ++h) nr[h] = nu * rho[h];
  C<T, K> xm = prev1(x), xp = next1(x);
  C<T, K> nm = prev1(nr), np = next1(nr);
  C<T, K> um = prev1(u), up = next1(u);
  C<T, K> vm = prev1(v), vp = next1(v);
  C<T, K> wm = prev1(w), wp = next1(w);
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const int c = cell_of(K, h);
    T n = nr[h], rh = rho[h];
    T dn, d1, d2;
    derivs<T>(xm[h], x[h], xp[h], nm[h], n, np[h], dn, d2);
    derivs<T>(xm[h], x[h], xp[h], wm[h], w[h], wp[h], d1, d2);
    T ddw = T(2.0) * n * d2 + T(2.0) * dn * d1;
    derivs<T>(xm[h], x[h], xp[h], um[h], u[h], up[h], d1, d2);
    T ddu = n * d2 + dn * d1;
    derivs<T>(xm[h], x[h], xp[h], vm[h], v[h], vp[h], d1, d2);
    T ddv = n * d2 + dn * d1;
    ddw = w[h] + dt * ddw / rh;
    ddu = u[h] + dt * ddu / rh;
    ddv = v[h] + dt * ddv / rh;
    if (c >= kNpad && c < lim - kNpad) {
      w[h] = ddw;
      u[h] = ddu;
      v[h] = ddv;
    }
  }
}

}  // namespace pn
