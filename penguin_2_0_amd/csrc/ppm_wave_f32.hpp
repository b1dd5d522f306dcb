// ppm_wave_f32.hpp -- the PPMLR window update in perturbation form, FP32 bulk (2-D
// isothermal disk).
//
// Method and stage order are those of ppm_wave.hpp (ppmlr.cu:2-56); one wavefront owns a
// window, K cells per lane.  The arithmetic is arranged so that FP32 keeps the accuracy
// that matters for a disk whose state is a smooth background plus small perturbations:
//   * density, pressure and angular momentum are carried as deviations from per-ring
//     backgrounds (rho_b, p_b = c_s^2 rho_b, L_b = r v_bg).  A parabola is built in each
//     cell's own frame: neighbour differences are (deviation difference) + (background
//     difference, an FP64 value rounded once), so no FP32 number holds a large value next
//     to a small change of it;
//   * the Riemann solve works on relative jumps eps = p / p_b(left cell) - 1 with accurate
//     log1p / expm1, and returns p* as such a relative jump;
//   * the radial momentum balance -- star gravity, centrifugal l^2/r^3 and the pressure
//     jump across the cell, which nearly cancel -- is split into its background residual
//     (evaluated once per ring in FP64) and terms that are each small, so every FP32
//     rounding is relative to a small quantity;
//   * coordinates in offset form (edge displacements dlt from the Eulerian edges);
//   * mass and angular momentum remapped in offset form, the background part of each flux
//     difference taken exactly (compensated flux differencing);
//   * the planet's force from the half-angle form R^2 = (r-r_p)^2 + 4 r r_p sin^2(dphi/2),
//     half-angle sin / cos from a per-column FP64 table rotated by short Taylor series.
// The azimuthal sweep advects v_phi - v_bg (an exact Galilean shift of the sweep) and moves
// its grid at the per-ring residual rate (v_bg - res_v)/r - FrRot; that uniform part of the
// grid motion is kept apart from the velocity-dependent part so the cell volumes and the
// density fluxes never difference it.
#pragma once

#include "pn_f32.hpp"
#include "ppm_wave.hpp"

namespace pn {

// A raw buffer view of a device array: loads and stores take 32-bit byte offsets, so an
// access costs no 64-bit address arithmetic; offsets at or past nbytes read 0 and drop
// stores (the hardware range check).  Word 3 of the descriptor: raw 32-bit data (gfx9).
// the wide buffer loads return integer vectors: take them whole and bit-cast the vector
typedef int i32x2_t __attribute__((ext_vector_type(2)));
typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

struct Buf {
  __amdgpu_buffer_rsrc_t r;
  Buf() = default;
  __device__ __forceinline__ Buf(const void* p, unsigned nbytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)nbytes, 0x00020000)) {}
  __device__ __forceinline__ float f(unsigned off) const {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
  }
  __device__ __forceinline__ float2 f2(unsigned off) const {
    const f32x2_t v = __builtin_bit_cast(f32x2_t, (i32x2_t)__builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
    return make_float2(v.x, v.y);
  }
  __device__ __forceinline__ void put(float v, unsigned off) const {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, 0);
  }
};

// m, x: the planet's mass and distance as the force forms take them (about the star, or in
// the barycentric frame m/(1+m) and x/(1+m)); inv_x2: the indirect term's 1/x^2 (0 in the
// barycentric frame); stx, stm: the star's distance from the barycentre and its mass 1/(1+m)
// (barycentric frame only, bary = 1)
struct PlanetF {
  float m, x, rs2, inv_x2, vy, stx, stm;
  int bary;
};

// The star's pull in the barycentric frame (bary_flag, forces.cu:28-40, 116-126): the star
// sits at stx opposite the planet, so with cf = cos(dphi) (times sin theta in 3-D)
// Rs^2 = r^2 + A, A = stx (stx + 2 r cf).  star_ri3 is 1/Rs^3; star_fr_rest the radial pull
// beyond the -1/r^2 the backgrounds balance and the monopole m/(1+m)/r^2 the callers add,
//   -stm (r + stx cf)/Rs^3 + stm/r^2 = stm (A (Rs^2 + Rs r + r^2) / ((Rs + r) r^2) - stx cf) / Rs^3,
// every term of the order of m (no cancellation against 1/r^2).
__device__ __forceinline__ float star_ri3(const PlanetF& p, float r, float cf) {
  const float ri = __frsqrt_rn(r * r + p.stx * (p.stx + 2.0f * r * cf));
  return ri * ri * ri;
}
__device__ __forceinline__ float star_fr_rest(const PlanetF& p, float r, float cf) {
  const float A = p.stx * (p.stx + 2.0f * r * cf);
  const float Rs2 = r * r + A;
  const float ri = __frsqrt_rn(Rs2), Rs = Rs2 * ri;
  return p.stm * (ri * ri * ri) * (A * (Rs2 + Rs * r + r * r) / ((Rs + r) * r * r) - p.stx * cf);
}

// sin / cos of (A + e)/2 from sin / cos of A/2 and a small e (|e| < ~0.5)
__device__ __forceinline__ void rot_half(float s, float c, float e, float& so, float& co) {
  const float h = 0.5f * e, h2 = h * h;
  const float sh = h * (1.0f - h2 * (1.0f / 6.0f) * (1.0f - h2 * (1.0f / 20.0f) * (1.0f - h2 * (1.0f / 42.0f))));
  const float ch = 1.0f - h2 * 0.5f * (1.0f - h2 * (1.0f / 12.0f) * (1.0f - h2 * (1.0f / 30.0f)));
  so = s * ch + c * sh;
  co = c * ch - s * sh;
}

// planet gravity with the indirect term (forces.cu:112-159) from the half angle to the planet
template <bool BARY = false>
__device__ __forceinline__ float planet_fr(const PlanetF& p, float rad, float s, float c) {
  const float s2 = s * s;
  const float cosd = 1.0f - 2.0f * s2;
  const float dr = rad - p.x;
  const float R2 = dr * dr + 4.0f * rad * p.x * s2 + p.rs2;
  const float ri = __frsqrt_rn(R2);
  float f = -p.m * (dr + 2.0f * p.x * s2) * (ri * ri * ri) - p.m * cosd * p.inv_x2;
  if constexpr (BARY) if (p.bary) f += p.m / (rad * rad) + star_fr_rest(p, rad, cosd);
  return f;
}
template <bool BARY = false>
__device__ __forceinline__ float planet_fphi(const PlanetF& p, float rad, float s, float c) {
  const float s2 = s * s;
  const float sind = 2.0f * s * c;
  const float dr = rad - p.x;
  const float R2 = dr * dr + 4.0f * rad * p.x * s2 + p.rs2;
  const float ri = __frsqrt_rn(R2);
  float f = -p.m * p.x * sind * (ri * ri * ri) + p.m * sind * p.inv_x2;
  if constexpr (BARY) if (p.bary) f += p.stm * p.stx * sind * star_ri3(p, rad, 1.0f - 2.0f * s2);
  return f;
}

// spherical forms (forces.cu:23-110) for a planet in the midplane and a cell at (r, theta),
// st = sin theta, oms = 1 - sin theta: D^2 = r^2 + r_p^2 - 2 r r_p cos(dphi) sin(theta)
// = (r - r_p)^2 + 2 r r_p k with k = oms + 2 st sin^2(dphi/2) = 1 - cos(dphi) st.  Here p.m is
// the reduced mass m/(1+m) of these forms; the star's -1/(1+m)/r^2 = -1/r^2 + p.m/r^2 is
// left to the caller.
template <bool BARY = false>
__device__ __forceinline__ float planet_fr_sph(const PlanetF& p, float r, float st, float oms, float s, float c) {
  const float s2 = s * s;
  const float cosd = 1.0f - 2.0f * s2;
  const float k = oms + 2.0f * st * s2;
  const float dr = r - p.x;
  const float ri = __frsqrt_rn(dr * dr + 2.0f * r * p.x * k + p.rs2);
  float f = -p.m * (dr + p.x * k) * (ri * ri * ri) - p.m * st * cosd * p.inv_x2;
  if constexpr (BARY) if (p.bary) f += star_fr_rest(p, r, st * cosd);
  return f;
}
template <bool BARY = false>
__device__ __forceinline__ float planet_fphi_sph(const PlanetF& p, float r, float st, float oms, float s, float c) {
  const float s2 = s * s;
  const float sind = 2.0f * s * c;
  const float k = oms + 2.0f * st * s2;
  const float dr = r - p.x;
  const float ri = __frsqrt_rn(dr * dr + 2.0f * r * p.x * k + p.rs2);
  float f = -p.m * p.x * sind * (ri * ri * ri) + p.m * st * sind * p.inv_x2;
  if constexpr (BARY) if (p.bary) f += p.stm * p.stx * sind * star_ri3(p, r, st * (1.0f - 2.0f * s2));
  return f;
}
template <bool BARY = false>
__device__ __forceinline__ float planet_fth_sph(const PlanetF& p, float r, float st, float ct, float oms, float s,
                                                float c) {
  const float s2 = s * s;
  const float cosd = 1.0f - 2.0f * s2;
  const float k = oms + 2.0f * st * s2;
  const float dr = r - p.x;
  const float ri = __frsqrt_rn(dr * dr + 2.0f * r * p.x * k + p.rs2);
  float f = p.m * p.x * cosd * ct * (ri * ri * ri) - p.m * cosd * ct / st * p.inv_x2;
  if constexpr (BARY) if (p.bary) f -= p.stm * p.stx * cosd * ct * star_ri3(p, r, st * cosd);
  return f;
}

// true in every lane when the condition holds in any lane of the wavefront (wave-uniform)
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// log(1 + x): 2 atanh(x / (2 + x)) as an odd series for |x| < 1/4, else the library log.
// Deviations that large are rare (next to a planet), so the library branch is taken only by
// wavefronts that hold one; elsewhere it costs a compare and a scalar branch.
__device__ __forceinline__ float log1p_acc(float x) {
  const float z = x / (2.0f + x), z2 = z * z;
  float s = 2.0f * z * (1.0f + z2 * (1.0f / 3.0f + z2 * (1.0f / 5.0f + z2 * (1.0f / 7.0f + z2 * (1.0f / 9.0f)))));
  if (wave_any(!(fabsf(x) < 0.25f))) s = fabsf(x) < 0.25f ? s : __logf(1.0f + x);
  return s;
}
// exp(y) - 1: Taylor series to y^7 for |y| < 1/4, else the library exp (branch as above)
__device__ __forceinline__ float expm1_acc(float y) {
  float s =
      y * (1.0f + y * (0.5f + y * (1.0f / 6.0f + y * (1.0f / 24.0f + y * (1.0f / 120.0f + y * (1.0f / 720.0f + y * (1.0f / 5040.0f)))))));
  if (wave_any(!(fabsf(y) < 0.25f))) s = fabsf(y) < 0.25f ? s : __expf(y) - 1.0f;
  return s;
}

// paraset of a uniform grid (CW84 eq. 1.6 with equal widths): constants
template <int K>
__device__ __forceinline__ Para<float, K> paraset_uniform() {
  Para<float, K> P;
  P.c0 = splat<float, K>(0.5f);
  P.c1 = splat<float, K>(-1.0f / 6.0f);
  P.c2 = splat<float, K>(1.0f / 6.0f);
  P.c3 = splat<float, K>(0.5f);
  P.c4 = splat<float, K>(0.5f);
  return P;
}

// parabola of X = X_b + x in each cell's frame (parabola.cu:70-136), in relative form: val = x
// (the cell's deviation), dbg = X_b(c+1) - X_b(c) (BG = false: a uniform background, dbg
// unused).  The edges are carried as offsets from the cell's own value (the right edge
// c0 da_r + c1 slope(c+1) + c2 slope, the left edge the left neighbour's right edge in this
// frame), so the flattening, the extremum and overshoot tests, the jump and the curvature
// never subtract two values of the cell's size; the algebra is the reference's.  Outputs:
// al = left edge - x, da, a6, and the jumps dl = X(c) - X(c-1), dr = X(c+1) - X(c) in this
// cell's frame.  A cell outside [nmin+2, nmaxp-2) returns the plain parabola's zero edge
// (al = -x, da = a6 = 0).
template <int K>
struct ParabR {
  C<float, K> al, da, a6, dl, dr;
};

template <int K, int SL = 64, bool BG = true>
__device__ __forceinline__ ParabR<K> parabola_rel(int nmin, int nmaxp, const C<float, K>& flat,
                                                  const C<float, K>& val, const C<float, K>& dbg,
                                                  const Para<float, K>& P) {
  C<float, K> vm = prev1(val), vp = next1(val), dbm;
  if constexpr (BG) dbm = prev1(dbg);
  ParabR<K> o;
  C<float, K> slope;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const float a = val[h];
    o.dr[h] = BG ? (vp[h] - a) + dbg[h] : vp[h] - a;
    o.dl[h] = BG ? (a - vm[h]) + dbm[h] : a - vm[h];
    slope[h] = limited_slope<float>(P.c3[h] * o.dr[h] + P.c4[h] * o.dl[h], o.dl[h], o.dr[h]);
  }
  C<float, K> slp = next1(slope);
  C<float, K> ar;   // right edge - x
#pragma unroll
  for (int h = 0; h < K; ++h) ar[h] = P.c0[h] * o.dr[h] + P.c1[h] * slp[h] + P.c2[h] * slope[h];
  C<float, K> arm = prev1(ar);
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const int c = cell_in<SL>(K, h);
    const float g = 1.0f - flat[h];
    float r = g * ar[h];                 // flattened right edge - x
    float l = g * (arm[h] - o.dl[h]);    // flattened left edge - x
    const float d = r - l;
    const float t2 = d * d;
    const float t1 = -3.0f * d * (r + l);   // (ar - la) 6 (x - (ar + la)/2)
    if (r * l >= 0.0f) {                 // (ar - x)(x - la) <= 0: an extremum
      r = 0.0f;
      l = 0.0f;
    }
    if (t2 < t1) l = -2.0f * r;          // la = 3 x - 2 ar
    if (t2 < -t1) r = -2.0f * l;         // ar = 3 x - 2 la
    const bool in = (c >= nmin + 2 && c < nmaxp - 2);
    o.da[h] = in ? r - l : 0.0f;
    o.a6[h] = in ? -3.0f * (r + l) : 0.0f;
    o.al[h] = in ? l : -val[h];
  }
  return o;
}

// sweep geometries of the FP32 engine: radial (cylindrical, 2-D), azimuthal, radial
// (spherical, 3-D), polar angle theta (3-D)
enum SweepGeo { kGeoRcyl = 0, kGeoPhi = 1, kGeoRsph = 2, kGeoTheta = 3 };

template <int K>
struct WinP {
  // carried state: deviations from the cell's backgrounds
  C<float, K> r;   // rho - rho_b
  C<float, K> u;   // sweep-normal velocity: v_r (radial), v_phi - v_bg (azimuthal), v_theta (theta)
  C<float, K> v;   // radial: l - L_b (l = r_cyl v_phi); azimuthal and theta: v_r
  C<float, K> w;   // 3-D only -- spherical radial: r v_theta; azimuthal: v_theta; theta: l - L_b
  C<float, K> q;   // adiabatic (EOS 3, 2-D): p / rho - q_b
  // backgrounds (azimuthal sweep: uniform in the window, differences zero); dLb is the
  // difference of L_b along the sweep
  C<float, K> rb, drb, pb, dpb, Lb, dLb, cs, cs2;
  C<float, K> qb, dqb;   // EOS 3: q_b = p_b / rho_b and its difference along the sweep
  // geometry; sth / oms: sin theta and 1 - sin theta of the cell centre (3-D planet terms);
  // vrot (azimuthal sweep) or rs / rcs (radial and theta sweeps: sin / cos of the planet's
  // drift against the slot in half a step, halved) for the planet's azimuth over the step
  C<float, K> rad, xa0, dx0, dvol0, vrot, ts, tc, sth, oms, rs, rcs;
};

// the per-slot table entries of this lane's K cells, entry-major (entry e of slot s at
// float e * stride + s): one vector buffer load per entry, the lane's byte offset in a
// register and the entry's offset e * stride * 4 in a scalar one
template <int K>
struct TabLane {
  Buf b;           // the table
  unsigned voff;   // byte offset of the slot of this lane's first cell
  int stride;
  __device__ __forceinline__ C<float, K> get(int e) const {
    const int soff = e * stride * 4;
    C<float, K> o;
    if constexpr (K == 2) {
      const f32x2_t v =
          __builtin_bit_cast(f32x2_t, (i32x2_t)__builtin_amdgcn_raw_buffer_load_b64(b.r, (int)voff, soff, 0));
      o[0] = v.x;
      o[1] = v.y;
    } else if constexpr (K == 4) {
      const f32x4_t v =
          __builtin_bit_cast(f32x4_t, (i32x4_t)__builtin_amdgcn_raw_buffer_load_b128(b.r, (int)voff, soff, 0));
      o[0] = v.x;
      o[1] = v.y;
      o[2] = v.z;
      o[3] = v.w;
    } else {
      // K = 3: three dword loads, which the compiler merges into one buffer_load_dwordx3
      // (the 96-bit load builtin itself is not used: this compiler fetches only its first
      // dword)
#pragma unroll
      for (int h = 0; h < K; ++h)
        o[h] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(b.r, (int)(voff + 4 * h), soff, 0));
    }
    return o;
  }
};

// the same for cells `step` slots apart (the theta sweep: one theta row of slots); c0 = this
// lane's first cell of the window, reads past cell cmax (the window's last) repeat it
template <int K>
struct TabLaneS {
  const float* __restrict__ p;
  int stride, base, step, c0, cmax;
  __device__ __forceinline__ C<float, K> get(int e) const {
    const float* q = p + (long)e * stride + base;
    C<float, K> o;
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const int c = c0 + h;
      o[h] = q[(long)step * (c < cmax ? c : cmax)];
    }
    return o;
  }
};

// the half angle (s, c) turned by the slot's half-step rotation (rs, rc), and by twice that
// (a full step)
__device__ __forceinline__ void rot_by(float s, float c, float rs, float rc, float& so, float& co) {
  so = s * rc + c * rs;
  co = c * rc - s * rs;
}
__device__ __forceinline__ void rot_by2(float s, float c, float rs, float rc, float& so, float& co) {
  rot_by(s, c, 2.0f * rs * rc, 1.0f - 2.0f * rs * rs, so, co);
}

// cos(a + d) - cos(a) from sin a, cos a and a small d (|d| < ~0.3), without cancellation
__device__ __forceinline__ float dcos(float sa, float ca, float d) {
  const float d2 = d * d;
  const float sd = d * (1.0f - d2 * (1.0f / 6.0f) * (1.0f - d2 * (1.0f / 20.0f)));
  const float omc = 0.5f * d2 * (1.0f - d2 * (1.0f / 12.0f) * (1.0f - d2 * (1.0f / 30.0f)));
  return -sa * sd - ca * omc;
}

// sin / cos / 1 - sin of (theta + e) from those of theta and a small e
__device__ __forceinline__ void rot_theta(float st, float ct, float oms, float e, float& so, float& co,
                                          float& omso) {
  const float e2 = e * e;
  const float se = e * (1.0f - e2 * (1.0f / 6.0f) * (1.0f - e2 * (1.0f / 20.0f)));
  const float omc = 0.5f * e2 * (1.0f - e2 * (1.0f / 12.0f) * (1.0f - e2 * (1.0f / 30.0f)));
  so = st * (1.0f - omc) + ct * se;
  co = ct * (1.0f - omc) - st * se;
  omso = oms + st * omc - ct * se;
}

// The window update.  G: sweep geometry; NDIM: 2 or 3 (the azimuthal sweep carries w and uses
// the spherical planet terms in 3-D); phi_rows: the azimuthal sweep's W.rs / W.rcs hold the
// rotation from a column's left edge over half a step (uniform azimuthal grid).  tx: the per-slot table of this lane's cells (TabLane
// or TabLaneS; radial and theta sweeps read their balance terms where they are used).  P0: paraset of the
// Eulerian grid (its values outside the window's paraset ranges never reach a parabola the
// sweep keeps).  wres: the azimuthal grid's uniform angular rate.  rsph: the spherical
// radius of an azimuthal 3-D ring.  On return cells [6, nmax-6) hold the update.
template <int K, int G, int NDIM, class Tab, bool ADIA = false, int SL = 64, bool BARY = false>
__device__ __forceinline__ void ppmlr_pert(WinP<K>& W, const Tab& tx, const Para<float, K>& P0, int nmax,
                                           float dt, float wres, const PlanetF& pl, bool planet, bool flat_on,
                                           float sqrt_gam, float rsph = 0.0f, bool phi_rows = false,
                                           float gam = 1.0f) {
  constexpr bool RAD = (G == kGeoRcyl || G == kGeoRsph);
  static_assert(!ADIA || (NDIM == 2 && (G == kGeoRcyl || G == kGeoPhi)), "EOS 3 is built for 2-D disks");
  const float hdt = 0.5f * dt;
  const float fourthd = 1.3333333333333333f;
  const float smallr = 1.0e-15f;
  // EOS 3 carries q = p / rho as its deviation dq from q_b; the pressure deviation is
  // p - p_b = rho_b dq + q_b drho + drho dq (p_b = rho_b q_b)
  C<float, K> dpc;
  if constexpr (ADIA) {
#pragma unroll
    for (int h = 0; h < K; ++h) dpc[h] = W.rb[h] * W.q[h] + W.qb[h] * W.r[h] + W.r[h] * W.q[h];
  }

  // ---- flatten (flatten.cu:7-40), on absolute pressure (relative-jump thresholds)
  C<float, K> flat = splat<float, K>(0.0f);
  if (flat_on) {
    C<float, K> pabs;
#pragma unroll
    for (int h = 0; h < K; ++h) {
      if constexpr (ADIA) pabs[h] = W.pb[h] + dpc[h];
      else pabs[h] = W.pb[h] + W.r[h] * W.cs2[h];
    }
    C<float, K> pm1 = prev1(pabs), pp1 = next1(pabs), pm2 = prev2(pabs), pp2 = next2(pabs);
    C<float, K> um1 = prev1(W.u), up1 = next1(W.u);
    C<float, K> steep;
#pragma unroll
    for (int h = 0; h < K; ++h) {
      // evaluated in every lane and selected (no divergent branch around it)
      const int c = cell_in<SL>(K, h);
      const float dp1 = pp1[h] - pm1[h];
      float dp2 = pp2[h] - pm2[h];
      dp2 = (fabsf(dp2) < 1.0e-15f) ? 1.0e-15f : dp2;
      const float ddp = fabsf(dp1) / fminf(pp1[h], pm1[h]);
      const bool shock = (c >= 2 && c < nmax - 2) && ddp > 0.33f && !(um1[h] < up1[h]);
      const float st = fmaxf(0.0f, (dp1 / dp2 - 0.75f) * 5.0f);
      steep[h] = shock ? st : 0.0f;
    }
    C<float, K> sm = prev1(steep), sp = next1(steep);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const int c = cell_in<SL>(K, h);
      const float fl = fminf(0.5f, fmaxf(sm[h], fmaxf(steep[h], sp[h])));
      flat[h] = (c >= 2 && c < nmax - 2) ? fl : 0.0f;
    }
  }

  // ---- states (states.cu:3-73)
  C<float, K> fhalf;   // hdt * force at the left edge
  if constexpr (RAD) {
    C<float, K> vm = prev1(W.v);
    const C<float, K> ffb = tx.get(kTabFface), Lf = tx.get(kTabLf), ixa3 = tx.get(kTabIxa3);
    C<float, K> wm, ixa3w;
    if constexpr (G == kGeoRsph) {
      wm = prev1(W.w);
      ixa3w = tx.get(kTabIxa3w);
    }
#pragma unroll
    for (int h = 0; h < K; ++h) {
      // l at the edge = L_f + dlf; f = f_face + dlf (2 L_f + dlf) / (xa0^3 sin^2) (+ w^2 / xa0^3)
      const float dlf = 0.5f * (W.v[h] + vm[h]);
      float f = ffb[h] + dlf * (2.0f * Lf[h] + dlf) * ixa3[h];
      if constexpr (G == kGeoRsph) {
        const float wf = 0.5f * (W.w[h] + wm[h]);
        f += wf * wf * ixa3w[h];
      }
      if (planet) {
        float s, c;
        rot_by(W.ts[h], W.tc[h], W.rs[h], W.rcs[h], s, c);
        if constexpr (G == kGeoRcyl) {
          f += planet_fr<BARY>(pl, W.xa0[h], s, c);
        } else {
          const float ix = 1.0f / W.xa0[h];
          f += pl.m * ix * ix + planet_fr_sph<BARY>(pl, W.xa0[h], W.sth[h], W.oms[h], s, c);
        }
      }
      fhalf[h] = hdt * f;
    }
  } else if constexpr (G == kGeoPhi) {
#pragma unroll
    for (int h = 0; h < K; ++h) {
      float f = 0.0f;
      if (planet) {
        float s, c;
        // from the left edge of the column to the planet half a step on: the ring's own
        // rotation rows on a uniform grid, else the series
        if (phi_rows) rot_by(W.ts[h], W.tc[h], W.rs[h], W.rcs[h], s, c);
        else rot_half(W.ts[h], W.tc[h], -0.5f * W.dx0[h] + hdt * (W.vrot[h] - pl.vy), s, c);
        if constexpr (NDIM == 2) f = planet_fphi<BARY>(pl, W.rad[h], s, c);
        else f = planet_fphi_sph<BARY>(pl, rsph, W.sth[h], W.oms[h], s, c);
      }
      fhalf[h] = hdt * f;
    }
  } else {
    // theta edge: f = f_face + dlf (2 L_f + dlf) cos / (r sin)^3 of the edge (+ planet)
    C<float, K> wm = prev1(W.w);
    const C<float, K> ffb = tx.get(kTabFfaceZ), Lf = tx.get(kTabLfZ), Gf = tx.get(kTabGfZ);
    C<float, K> szl, czl;
    if (planet) {
      szl = tx.get(kTabSzl);
      czl = tx.get(kTabCzl);
    }
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const float dlf = 0.5f * (W.w[h] + wm[h]);
      float f = ffb[h] + dlf * (2.0f * Lf[h] + dlf) * Gf[h];
      if (planet) {
        float s, c;
        rot_by(W.ts[h], W.tc[h], W.rs[h], W.rcs[h], s, c);
        f += planet_fth_sph<BARY>(pl, W.rad[h], szl[h], czl[h], 1.0f - szl[h], s, c);
      }
      fhalf[h] = hdt * f;
    }
  }
  C<float, K> csn, csfac;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    // sqrt(gam p / rho) = sqrt(gam) c_s for the isothermal closure, sqrt(gam q) for EOS 3
    float s;
    if constexpr (ADIA) s = hdt * sqrtf(gam * (W.qb[h] + W.q[h])) / W.dx0[h];
    else s = hdt * sqrt_gam * W.cs[h] / W.dx0[h];
    if constexpr (G == kGeoPhi || G == kGeoTheta) s = s / W.rad[h];
    csn[h] = s;
    csfac[h] = 1.0f - fourthd * s;
  }
  // edge averages of the traced parabolae (states.cu:45-68), in the cell's frame: the
  // cell's value x plus the average's offset from it
  auto edges = [&](const ParabR<K>& q, const C<float, K>& x, C<float, K>& re, C<float, K>& le) {
#pragma unroll
    for (int h = 0; h < K; ++h) {
      re[h] = x[h] + (q.al[h] + q.da[h] - csn[h] * (q.da[h] - csfac[h] * q.a6[h]));
      le[h] = x[h] + (q.al[h] + csn[h] * (q.da[h] + csfac[h] * q.a6[h]));
    }
  };
  // interface c: left state = right edge of cell c-1, right state = left edge of cell c, as
  // relative jumps against the backgrounds of cell c-1
  C<float, K> epl, epr, erl, err, ul, ur;
  if constexpr (G != kGeoPhi || ADIA) {
    C<float, K> pp, re, le;
#pragma unroll
    for (int h = 0; h < K; ++h) {
      if constexpr (ADIA) pp[h] = dpc[h];
      else pp[h] = W.r[h] * W.cs2[h];
    }
    // along a ring p_b is one constant
    edges(parabola_rel<K, SL, G != kGeoPhi>(0, nmax, flat, pp, W.dpb, P0), pp, re, le);
    C<float, K> q;
#pragma unroll
    for (int h = 0; h < K; ++h) q[h] = re[h] / W.pb[h];
    C<float, K> qm = prev1(q), pbm = prev1(W.pb), dpbm = prev1(W.dpb);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      epl[h] = qm[h];
      epr[h] = (le[h] + dpbm[h]) / pbm[h];
    }
  }
  {
    C<float, K> re, le;
    edges(parabola_rel<K, SL, G != kGeoPhi>(0, nmax, flat, W.r, W.drb, P0), W.r, re, le);
    C<float, K> q;
#pragma unroll
    for (int h = 0; h < K; ++h) q[h] = re[h] / W.rb[h];
    C<float, K> qm = prev1(q), rbm = prev1(W.rb), drbm = prev1(W.drb);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      erl[h] = qm[h];
      err[h] = (le[h] + drbm[h]) / rbm[h];
    }
  }
  {
    C<float, K> re, le;
    edges(parabola_rel<K, SL, false>(0, nmax, flat, W.u, W.u, P0), W.u, re, le);
    C<float, K> rem = prev1(re);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      ul[h] = rem[h] + fhalf[h];
      ur[h] = le[h] + fhalf[h];
    }
  }

  // ---- riemann, isothermal two-rarefaction closed form (riemann.cu:46-53):
  //   t1 = (sqrt(pl/rl) + sqrt(pr/rr))/2, t2 = -(ur-ul)/(2 t1), p* = sqrt(pl pr) e^t2,
  //   u* = ul - t1 log(p*/pl); here with pl = p_b(1+epl), ..., p* = p_b(1+em)
  C<float, K> umid, em;
  if constexpr (ADIA) {
    // gamma-law interface (riemann.cu:55-71, EOS 2's solver): the two-rarefaction estimate
    // of p*, then u* from each side's shock or rarefaction curve.  Every quantity is taken
    // relative to the background of cell c-1 (p_b, rho_b, c0^2 = gam q_b): a_s = c_s / c0,
    // N / (2 c0) = (a_l + a_r)/2 - (gam-1)(ur-ul)/(4 c0), D / (2 c0 p_b^-z) =
    // (a_l (1+epl)^-z + a_r (1+epr)^-z)/2, p* / p_b = (N / D)^(1/z) with every factor formed
    // as 1 + (a small number)
    const float z = 0.5f * (gam - 1.0f) / gam, iz = 1.0f / z, gm1 = gam - 1.0f, gp1 = gam + 1.0f;
    const C<float, K> qbm = prev1(W.qb);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const float lpl = log1p_acc(epl[h]), lpr = log1p_acc(epr[h]);
      const float lrl = log1p_acc(erl[h]), lrr = log1p_acc(err[h]);
      const float c0 = sqrtf(gam * qbm[h]);
      const float hl = 0.5f * (lpl - lrl), hr = 0.5f * (lpr - lrr);   // log(a_l), log(a_r)
      const float al1 = expm1_acc(hl), ar1 = expm1_acc(hr);
      const float nm1 = 0.5f * (al1 + ar1) - 0.25f * gm1 * (ur[h] - ul[h]) / c0;
      const float dm1 = 0.5f * (expm1_acc(hl - z * lpl) + expm1_acc(hr - z * lpr));
      const float lm = iz * (log1p_acc(nm1) - log1p_acc(dm1));   // log(p* / p_b)
      const float e_ = expm1_acc(lm);
      float t1, t2;   // right and left wave terms of u* (riemann.cu:64-70)
      if (e_ > epr[h]) t1 = (e_ - epr[h]) * sqrtf(2.0f * qbm[h] / ((1.0f + err[h]) * (gp1 * (1.0f + e_) + gm1 * (1.0f + epr[h]))));
      else t1 = 2.0f * c0 * (1.0f + ar1) * expm1_acc(z * (lm - lpr)) / gm1;
      if (e_ > epl[h]) t2 = (e_ - epl[h]) * sqrtf(2.0f * qbm[h] / ((1.0f + erl[h]) * (gp1 * (1.0f + e_) + gm1 * (1.0f + epl[h]))));
      else t2 = 2.0f * c0 * (1.0f + al1) * expm1_acc(z * (lm - lpl)) / gm1;
      em[h] = e_;
      umid[h] = 0.5f * (ur[h] + ul[h]) + 0.5f * (t1 - t2);
    }
  } else if constexpr (G != kGeoPhi) {
    C<float, K> csm = prev1(W.cs);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const float lpl = log1p_acc(epl[h]), lpr = log1p_acc(epr[h]);
      // t1 only scales small differences (of velocity and of log pressure), so its plain
      // FP32 relative accuracy suffices: no logs of the density states
      const float t1 = 0.5f * csm[h] * (sqrtf((1.0f + epl[h]) / (1.0f + erl[h])) + sqrtf((1.0f + epr[h]) / (1.0f + err[h])));
      const float t2 = -0.5f * (ur[h] - ul[h]) / t1;
      em[h] = expm1_acc(0.5f * (lpl + lpr) + t2);
      umid[h] = ul[h] - t1 * (0.5f * (lpr - lpl) + t2);
    }
  } else {
    // along a ring c_s^2 is one constant, P = c_s^2 Sigma in every cell, and every step of
    // the reconstruction is homogeneous of degree one in the values: the pressure states are
    // c_s^2 times the density states, so eps_p = eps_rho and t1 = c_s exactly
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const float lrl = log1p_acc(erl[h]), lrr = log1p_acc(err[h]);
      const float t1 = W.cs[h];
      const float t2 = -0.5f * (ur[h] - ul[h]) / t1;
      em[h] = expm1_acc(0.5f * (lrl + lrr) + t2);
      umid[h] = ul[h] - t1 * (0.5f * (lrr - lrl) + t2);
    }
  }

  // ---- evolve (evolve.cu:3-166): edges as displacements from the Eulerian edges.  The
  // reference moves edges [3, nmax-2) and cells [3, nmax-3).
  C<float, K> smv, dlt;
  const float Wd = (G == kGeoPhi) ? dt * wres : 0.0f;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const int c = cell_in<SL>(K, h);
    const bool mv = (c >= 3 && c < nmax - 2);
    float s;
    if constexpr (RAD) s = dt * umid[h];
    else s = dt * umid[h] / W.rad[h];
    smv[h] = mv ? s : 0.0f;
    dlt[h] = mv ? s + Wd : 0.0f;
  }
  C<float, K> Dl, dltx, sx;
  {
    C<float, K> sp = next1(smv);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const int c = cell_in<SL>(K, h);
      const bool in = (c >= 3 && c < nmax - 3);
      Dl[h] = in ? sp[h] - smv[h] : 0.0f;   // dxl - dx0, free of the uniform motion
      dltx[h] = in ? dlt[h] : 0.0f;          // the left edge of a moved cell
      sx[h] = in ? smv[h] : 0.0f;
    }
  }
  C<float, K> rL;   // Lagrangian density deviation
  C<float, K> szl, czl;   // theta sweep: sin / cos of the left edge
  if constexpr (G == kGeoTheta) {
    szl = tx.get(kTabSzl);
    czl = tx.get(kTabCzl);
  }
  {
    C<float, K> emp = next1(em);
    if constexpr (RAD) {
      C<float, K> pbm = prev1(W.pb), dltxp = next1(dltx), xa0p = next1(W.xa0);
      const C<float, K> tixc = tx.get(kTabIxc), res_bg = tx.get(kTabResBg), abg = tx.get(kTabAbg),
                        dpbm = tx.get(kTabDpbm);
      C<float, K> is2;
      if constexpr (G == kGeoRsph) is2 = tx.get(kTabIs2);
      // EOS 3: the pressure work through the left edge, (area of the moved edge) u* p*
      C<float, K> Fl, Flp;
      if constexpr (ADIA) {
#pragma unroll
        for (int h = 0; h < K; ++h) Fl[h] = (W.xa0[h] + 0.5f * dltx[h]) * umid[h] * (pbm[h] + pbm[h] * em[h]);
        Flp = next1(Fl);
      }
#pragma unroll
      for (int h = 0; h < K; ++h) {
        const int c = cell_in<SL>(K, h);
        const bool in = (c >= 3 && c < nmax - 3);
        const float ixc = tixc[h];
        const float ixc3 = ixc * ixc * ixc;
        const float rho = W.rb[h] + W.r[h];
        const float dl = W.v[h];
        // force at the Eulerian centre (f1) and at the moved centre (f2) less the
        // background's: gravity -1/r^2 and centrifugal l^2/r^3 (spherical: l^2/(r^3 sin^2)
        // + w^2/r^3) with l = L_b + dl
        const float il = (G == kGeoRsph) ? ixc3 * is2[h] : ixc3;
        float q1 = dl * (2.0f * W.Lb[h] + dl) * il;
        if constexpr (G == kGeoRsph) q1 += W.w[h] * W.w[h] * ixc3;
        const float dxc = dltx[h] + 0.5f * Dl[h];             // moved centre - centre
        const float e = dxc * ixc;
        const float g2 = -e * (2.0f - e * (3.0f - 4.0f * e));    // (1+e)^-2 - 1
        const float g3 = -e * (3.0f - e * (6.0f - 10.0f * e));   // (1+e)^-3 - 1
        const float B = W.Lb[h] * W.Lb[h] * il;
        float fdiff = -(ixc * ixc) * g2 + (B + q1) * g3;
        float f21 = fdiff;   // force at the moved centre - at the centre (EOS 3)
        if (planet) {
          float s, cc;
          rot_by2(W.ts[h], W.tc[h], W.rs[h], W.rcs[h], s, cc);
          if constexpr (G == kGeoRcyl) {
            const float p1 = planet_fr<BARY>(pl, W.rad[h], W.ts[h], W.tc[h]), p2 = planet_fr<BARY>(pl, W.rad[h] + dxc, s, cc);
            fdiff += p1 + p2;
            f21 += p2 - p1;
          } else {
            // the star pulls with 1/(1+m): -1/r^2 + plm/r^2 at both centres
            fdiff += pl.m * ixc * ixc * (2.0f + g2) + planet_fr_sph<BARY>(pl, W.rad[h], W.sth[h], W.oms[h], W.ts[h], W.tc[h]) +
                     planet_fr_sph<BARY>(pl, W.rad[h] + dxc, W.sth[h], W.oms[h], s, cc);
          }
        }
        // pressure: (p*(c+1) - p*(c)) area / (rho dvol0), area = V0 + dV, and
        //   p*(c+1) - p*(c) = dp_bm + dp,  dp = p_b(c) em(c+1) - p_b(c-1) em(c)
        const float dp = W.pb[h] * emp[h] - pbm[h] * em[h];
        float V0, dV;
        const float d = dltx[h], x0 = W.xa0[h], w0 = W.dx0[h], D = Dl[h];
        if constexpr (G == kGeoRcyl) {
          V0 = W.rad[h];
          dV = 0.25f * (d + dltxp[h]);
        } else {
          // area factor at an edge x0 moved by d: (x0 + d)^2 - d (x0 + d) + d^2/3 = x0^2 + d (x0 + d/3)
          const float dp_ = dltxp[h];
          V0 = 0.5f * (x0 * x0 + xa0p[h] * xa0p[h]);
          dV = 0.5f * (d * (x0 + d * (1.0f / 3.0f)) + dp_ * (xa0p[h] + dp_ * (1.0f / 3.0f)));
        }
        const float acc = res_bg[h] + q1 + 0.5f * fdiff + abg[h] * (W.r[h] / rho) -
                          (dpbm[h] * dV + dp * (V0 + dV)) / (rho * W.dvol0[h]);
        // Lagrangian volume change in offset form
        float ddv;
        if constexpr (G == kGeoRcyl) {
          // dvol = dxl (xa + dxl/2)
          ddv = w0 * (d + 0.5f * D) + D * (x0 + d + 0.5f * w0 + 0.5f * D);
        } else {
          // dvol = dxl (xa (xa + dxl) + dxl^2/3)
          const float a1 = x0 + d;
          ddv = w0 * d * (2.0f * x0 + d) + D * a1 * a1 + w0 * w0 * d + (2.0f * w0 * D + D * D) * a1 + w0 * w0 * D +
                w0 * D * D + D * D * D * (1.0f / 3.0f);
        }
        const float dvol = W.dvol0[h] + ddv;
        rL[h] = in ? W.r[h] - rho * ddv / dvol : W.r[h];   // (r dvol0 - rho_b ddv) / dvol
        if constexpr (ADIA) {
          // EOS 3 (the line evolve.cu:146 leaves commented out): e = eps + u^2/2 gains
          // -dt/dm (F_r - F_l) + dt (u f1 + u' f2)/2, eps' = e - u'^2/2.  With u' = u + du and
          // du = dt (f1 + f2)/2 - dt/dm dp A the body forces' work cancels against the kinetic
          // energy exactly, leaving
          //   eps' - eps = -dt/dm ((F_r - F_l) - (u + du/2) dp A) + dt du (f2 - f1) / 4
          const float du = dt * acc;
          const float dpA = dpbm[h] * V0 + (dpbm[h] * dV + dp * (V0 + dV));
          const float de = -dt / (rho * W.dvol0[h]) * ((Flp[h] - Fl[h]) - (W.u[h] + 0.5f * du) * dpA) +
                           0.25f * dt * du * f21;
          if (in) W.q[h] += (gam - 1.0f) * de;
        }
        if (in) W.u[h] = W.u[h] + dt * acc;
      }
    } else if constexpr (G == kGeoPhi) {
      // EOS 3: the pressure work through the left edge, u* p* (unit area factor)
      C<float, K> Fl, Flp;
      if constexpr (ADIA) {
#pragma unroll
        for (int h = 0; h < K; ++h) Fl[h] = umid[h] * (W.pb[h] + W.pb[h] * em[h]);
        Flp = next1(Fl);
      }
#pragma unroll
      for (int h = 0; h < K; ++h) {
        const int c = cell_in<SL>(K, h);
        const bool in = (c >= 3 && c < nmax - 3);
        float f = 0.0f, f21 = 0.0f;
        if (planet) {
          const float off = dltx[h] + 0.5f * Dl[h] + dt * (W.vrot[h] - pl.vy);
          float s, cc;
          rot_half(W.ts[h], W.tc[h], off, s, cc);
          if constexpr (NDIM == 2) {
            const float p1 = planet_fphi<BARY>(pl, W.rad[h], W.ts[h], W.tc[h]), p2 = planet_fphi<BARY>(pl, W.rad[h], s, cc);
            f = p1 + p2;
            f21 = p2 - p1;
          } else {
            f = planet_fphi_sph<BARY>(pl, rsph, W.sth[h], W.oms[h], W.ts[h], W.tc[h]) +
                planet_fphi_sph<BARY>(pl, rsph, W.sth[h], W.oms[h], s, cc);
          }
        }
        const float rho = W.rb[h] + W.r[h];
        const float dpm = W.pb[h] * (emp[h] - em[h]);   // p_b is uniform along the ring
        const float un = W.u[h] + (0.5f * dt * f - dt * dpm / (rho * W.dvol0[h]));
        const float ddv = Dl[h] * W.rad[h];
        const float dvol = W.dvol0[h] + ddv;
        rL[h] = in ? W.r[h] - rho * ddv / dvol : W.r[h];   // (r dvol0 - rho_b ddv) / dvol
        if constexpr (ADIA) {   // as in the radial sweep, with unit area factors
          const float du = un - W.u[h];
          const float de = -dt / (rho * W.dvol0[h]) * ((Flp[h] - Fl[h]) - (W.u[h] + 0.5f * du) * dpm) +
                           0.25f * dt * du * f21;
          if (in) W.q[h] += (gam - 1.0f) * de;
        }
        if (in) W.u[h] = un;
      }
    } else {
      // theta: centrifugal theta component L^2 G(theta), G = cos / (r sin)^3, minus the
      // background's; G at the moved centre by its Taylor series; the area factor at an edge
      // moved by d is (cos x0 - cos(x0 + d)) / d (sin x0 for d = 0)
      C<float, K> pbm = prev1(W.pb), dltxp = next1(dltx), szlp = next1(szl);
      const C<float, K> res_bg = tx.get(kTabResBgZ), abg = tx.get(kTabAbgZ), dpbm = tx.get(kTabDpbmZ),
                        Gc = tx.get(kTabGcZ), G1 = tx.get(kTabGc1Z), G2 = tx.get(kTabGc2Z),
                        szr = tx.get(kTabSzr), czr = tx.get(kTabCzr), cth = tx.get(kTabCth);
      C<float, K> vfd;   // area factor of the moved left edge minus sin x0
#pragma unroll
      for (int h = 0; h < K; ++h) {
        const float d = dltx[h], d2 = d * d;
        vfd[h] = szl[h] * d2 * (-1.0f / 6.0f + d2 * (1.0f / 120.0f)) + czl[h] * d * (0.5f - d2 * (1.0f / 24.0f));
      }
      C<float, K> vfdp = next1(vfd);
#pragma unroll
      for (int h = 0; h < K; ++h) {
        const int c = cell_in<SL>(K, h);
        const bool in = (c >= 3 && c < nmax - 3);
        const float rho = W.rb[h] + W.r[h];
        const float dl = W.w[h];
        const float qq = dl * (2.0f * W.Lb[h] + dl);
        const float L2 = W.Lb[h] * W.Lb[h] + qq;
        const float e = dltx[h] + 0.5f * Dl[h];   // angle of the moved centre - centre
        // 4 x (the mean of the forces at both centres - the background's): centrifugal
        // L^2 (G(theta + e) - G(theta)) and the planet at both centres, twice
        float fdiff = L2 * e * (2.0f * G1[h] + e * G2[h]);
        if (planet) {
          float s, cc, so, co, omso;
          rot_by2(W.ts[h], W.tc[h], W.rs[h], W.rcs[h], s, cc);
          rot_theta(W.sth[h], cth[h], W.oms[h], e, so, co, omso);
          fdiff += 2.0f * (planet_fth_sph<BARY>(pl, W.rad[h], W.sth[h], cth[h], W.oms[h], W.ts[h], W.tc[h]) +
                           planet_fth_sph<BARY>(pl, W.rad[h], so, co, omso, s, cc));
        }
        const float dp = W.pb[h] * emp[h] - pbm[h] * em[h];
        const float V0 = 0.5f * (szl[h] + szlp[h]);
        const float dV = 0.5f * (vfd[h] + vfdp[h]);
        const float acc = res_bg[h] + qq * Gc[h] + 0.25f * fdiff + abg[h] * (W.r[h] / rho) -
                          (dpbm[h] * dV + dp * (V0 + dV)) / (rho * W.dvol0[h]);
        // dvol = r (cos(x0 + d) - cos(x1 + d + D)) with x1 the right edge
        const float ddv = W.rad[h] * (dcos(szl[h], czl[h], dltx[h]) - dcos(szr[h], czr[h], dltx[h] + Dl[h]));
        const float dvol = W.dvol0[h] + ddv;
        rL[h] = in ? W.r[h] - rho * ddv / dvol : W.r[h];   // (r dvol0 - rho_b ddv) / dvol
        if (in) W.u[h] = W.u[h] + dt * acc;
      }
    }
  }

  // ---- remap (remap.cu:14-226)
  C<float, K> dxl;
#pragma unroll
  for (int h = 0; h < K; ++h) dxl[h] = W.dx0[h] + Dl[h];
  Para<float, K> P1 = paraset<float, K, SL>(3, nmax - 3, dxl);
  C<float, K> dxf, dxf2, lvol, delC;
  C<float, K> dxm = prev1(dxl);
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const float d = dltx[h];
    if constexpr (G == kGeoRcyl) lvol[h] = d * (W.xa0[h] + 0.5f * d);
    else if constexpr (G == kGeoRsph) lvol[h] = d * (W.xa0[h] * (W.xa0[h] + d) + d * d * (1.0f / 3.0f));
    else if constexpr (G == kGeoPhi) lvol[h] = d * W.rad[h];
    else lvol[h] = -W.rad[h] * dcos(szl[h], czl[h], d);
    delC[h] = (d >= 0.0f) ? 1.0f : 0.0f;
    float e;
    if (d >= 0.0f) {
      e = 0.5f * d / dxm[h];
      dxf2[h] = 1.0f - fourthd * e;
    } else {
      e = 0.5f * d / dxl[h];
      dxf2[h] = 1.0f + fourthd * e;
    }
    dxf[h] = e;
  }
  // average over the swept region of the left edge (remap.cu:59-90) minus the cell's own
  // value: from the left neighbour's parabola (its offset from X(c-1), then the jump
  // X(c) - X(c-1) in this frame) or from this cell's own
  auto face = [&](const ParabR<K>& Q, C<float, K>& out) {
    C<float, K> alm = prev1(Q.al), dam = prev1(Q.da), a6m = prev1(Q.a6);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      if (delC[h] != 0.0f) out[h] = (alm[h] + dam[h] - dxf[h] * (dam[h] - dxf2[h] * a6m[h])) - Q.dl[h];
      else out[h] = Q.al[h] - dxf[h] * (Q.da[h] + dxf2[h] * Q.a6[h]);
    }
  };
  C<float, K> invm, fr;
  {
    C<float, K> fv, frd;
    face(parabola_rel<K, SL, G != kGeoPhi>(3, nmax - 3, flat, rL, W.drb, P1), fv);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      fv[h] = rL[h] + fv[h];   // the face's deviation from rho_b of this cell
      if constexpr (G != kGeoPhi) {
        frd[h] = (W.rb[h] + fv[h]) * lvol[h];
        fr[h] = frd[h];
      } else {
        // rho_b times the uniform displacement is the same at every edge: it carries no
        // mass difference, so the density update leaves it out
        frd[h] = W.rb[h] * (W.rad[h] * sx[h]) + fv[h] * lvol[h];
        fr[h] = frd[h] + W.rb[h] * (W.rad[h] * (dltx[h] - sx[h]));
      }
    }
    C<float, K> frp = next1(frd);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const int c = cell_in<SL>(K, h);
      const float rn = W.r[h] + (frd[h] - frp[h]) / W.dvol0[h];
      invm[h] = 1.0f / (fmaxf(W.rb[h] + rn, smallr) * W.dvol0[h]);
      W.r[h] = (c >= 6 && c < nmax - 6) ? rn : rL[h];
    }
  }
  // X m_new = X m + F_left - F_right (remap.cu:93-131), taken in increment form: with
  // m_new - m = fr - fr(c+1) (the density update above),
  //   X_new = X + ((face_l - X) fr - (face_r - X) fr(c+1)) / m_new,
  // face_r the next cell's left face in this cell's frame.  Every rounding is then relative to
  // the increment, and a cell whose fluxes vanish keeps X exactly; the product form
  // (X m + ...) / m_new rounds X m / m_new away from X by up to an ulp of X, the same ulp
  // step after step in a steady flow, a drift of 1e-4 of X over 1000 steps.
  C<float, K> frp = next1(fr);
  auto remap_inc = [&](C<float, K>& X, const ParabR<K>& Q) {
    C<float, K> rl;
    face(Q, rl);
    C<float, K> rlp = next1(rl);
#pragma unroll
    for (int h = 0; h < K; ++h) {
      const int c = cell_in<SL>(K, h);
      // the right face relative to this cell's X: (face(c+1) - X(c+1)) + (X(c+1) + dbg - X)
      const float rr = rlp[h] + Q.dr[h];
      if (c >= 6 && c < nmax - 6) X[h] = X[h] + (rl[h] * fr[h] - rr * frp[h]) * invm[h];
    }
  };
  auto remap_plain = [&](C<float, K>& X) { remap_inc(X, parabola_rel<K, SL, false>(3, nmax - 3, flat, X, X, P1)); };
  // deviations from a background that varies along the sweep (l - L_b; EOS 3's q - q_b):
  // fluxes relative to the cell's own background; the right edge's face comes from the next
  // cell's frame, b(c+1) = b(c) + dbg
  auto remap_bg = [&](C<float, K>& X, const C<float, K>& dbg) {
    remap_inc(X, parabola_rel<K, SL, true>(3, nmax - 3, flat, X, dbg, P1));
  };
  remap_plain(W.u);
  if constexpr (RAD) remap_bg(W.v, W.dLb);
  else remap_plain(W.v);
  if constexpr (G == kGeoRsph || (G == kGeoPhi && NDIM == 3)) remap_plain(W.w);
  if constexpr (G == kGeoTheta) remap_bg(W.w, W.dLb);
  // EOS 3: q = p / rho, mass-weighted (remap.cu:171-176); q_b is uniform along a ring
  if constexpr (ADIA) {
    if constexpr (G == kGeoPhi) remap_plain(W.q);
    else remap_bg(W.q, W.dqb);
  }
}

// Explicit viscosity after the radial remap (device_viscosity_r, viscosity.cu:370-421) on
// the perturbation state: n = nu rho and Omega = l / R^2 (R = r_cyl) are background +
// deviation, and the quadratic-fit derivatives are linear in the values, so each derivative
// is the background's (from the table, FP64 rounded once) plus the deviation's (FP32, from
// the cell widths: the left edges of c-1, c, c+1 sit at 0, dx(c-1), dx(c-1) + dx(c)).
// l_new = l + dt R ddv / rho leaves L_b in place: dl += dt R ddv / rho.  SPH (3-D spherical
// r): also w = r v_theta, whose background is zero.
template <int K, int SL = 64, bool SPH = false>
__device__ __forceinline__ void visc_r_pert(WinP<K>& W, const TabLane<K>& tx, int lim, float dt) {
  const C<float, K> nu = tx.get(kTabNu), dnb = tx.get(kTabDnb), d1ob = tx.get(kTabD1Ob), d2ob = tx.get(kTabD2Ob),
                    irc = tx.get(kTabIrc), rcyl = tx.get(kTabRc);
  C<float, K> np_, om;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    np_[h] = nu[h] * W.r[h];
    om[h] = W.v[h] * (irc[h] * irc[h]);
  }
  C<float, K> wm = prev1(W.dx0);
  C<float, K> nm = prev1(np_), npp = next1(np_), um = prev1(W.u), up = next1(W.u), omm = prev1(om), omp = next1(om);
  C<float, K> tm, tp;
  if constexpr (SPH) {
    tm = prev1(W.w);
    tp = next1(W.w);
  }
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const int c = cell_in<SL>(K, h);
    const float x2 = wm[h], x3 = wm[h] + W.dx0[h], R = rcyl[h];
    const float rho = W.rb[h] + W.r[h], n = nu[h] * rho;
    float dn, d1, d2, e1, e2;
    derivs<float>(0.0f, x2, x3, nm[h], np_[h], npp[h], dn, d2);
    dn += dnb[h];
    derivs<float>(0.0f, x2, x3, um[h], W.u[h], up[h], d1, d2);
    const float ddu = 2.0f * n * d2 + 2.0f * dn * d1 + 2.0f * n * d1 / R - 2.0f * n * W.u[h] / R / R;
    derivs<float>(0.0f, x2, x3, omm[h], om[h], omp[h], e1, e2);
    e1 += d1ob[h];
    e2 += d2ob[h];
    const float ddv = R * n * e2 + R * dn * e1 + 3.0f * n * e1;
    float ddw = 0.0f;
    if constexpr (SPH) {
      derivs<float>(0.0f, x2, x3, tm[h], W.w[h], tp[h], d1, d2);
      ddw = n * d2 + dn * d1 + n * d1 / R;
    }
    if (c >= kNpad && c < lim - kNpad) {
      W.u[h] = W.u[h] + dt * ddu / rho;
      W.v[h] = W.v[h] + dt * R * ddv / rho;
      if constexpr (SPH) W.w[h] = W.w[h] + dt * ddw / rho;
    }
  }
}

// device_viscosity_p (viscosity.cu:423-472) on the azimuthal sweep's state: sweep u =
// v_phi - v_bg, sweep v = v_r and (3-D, W3) sweep w = v_theta.  nu and the backgrounds are
// uniform along the ring, so only deviations enter the derivatives.  The coordinates are
// arc lengths along the ring, the left edges times R = r_cyl (ppm.cu:277-278).
template <int K, bool W3 = false>
__device__ __forceinline__ void visc_p_pert(WinP<K>& W, int lim, float dt, float nu, float R) {
  C<float, K> np_, wx;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    np_[h] = nu * W.r[h];
    wx[h] = W.dx0[h] * R;
  }
  C<float, K> wm = prev1(wx);
  C<float, K> nm = prev1(np_), npp = next1(np_);
  C<float, K> vm = prev1(W.u), vp = next1(W.u), rm = prev1(W.v), rp = next1(W.v);
  C<float, K> tm, tp;
  if constexpr (W3) {
    tm = prev1(W.w);
    tp = next1(W.w);
  }
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const int c = cell_of(K, h);
    const float x2 = wm[h], x3 = wm[h] + wx[h];
    const float rho = W.rb[h] + W.r[h], n = nu * rho;
    float dn, d1, d2;
    derivs<float>(0.0f, x2, x3, nm[h], np_[h], npp[h], dn, d2);
    derivs<float>(0.0f, x2, x3, vm[h], W.u[h], vp[h], d1, d2);
    const float ddv = 2.0f * n * d2 + 2.0f * dn * d1;
    derivs<float>(0.0f, x2, x3, rm[h], W.v[h], rp[h], d1, d2);
    const float ddu = n * d2 + dn * d1;
    float ddw = 0.0f;
    if constexpr (W3) {
      derivs<float>(0.0f, x2, x3, tm[h], W.w[h], tp[h], d1, d2);
      ddw = n * d2 + dn * d1;
    }
    if (c >= kNpad && c < lim - kNpad) {
      W.u[h] = W.u[h] + dt * ddv / rho;
      W.v[h] = W.v[h] + dt * ddu / rho;
      if constexpr (W3) W.w[h] = W.w[h] + dt * ddw / rho;
    }
  }
}

// device_viscosity_z (viscosity.cu:474-518) on the theta sweep's state: sweep u = v_theta
// (along the pencil: 2 n d2 + 2 dn d1), sweep v = v_r and v_phi (n d2 + dn d1), with
// n = vis_nu rho and no curvature terms.  The coordinates are r times the theta left edges.
// rho_b and v_bg vary along theta: their derivatives come from the table (FP64 rounded once),
// the deviations' from the cell widths.  v_phi is carried as l - L_b = r_cyl (v_phi - v_bg):
// the update adds dt r_cyl ddv / rho to it.
template <int K, int SL = 64>
__device__ __forceinline__ void visc_z_pert(WinP<K>& W, const TabLane<K>& tx, int lim, float dt, float vis_nu) {
  const C<float, K> dnb = tx.get(kTabDnbZ), d1vb = tx.get(kTabD1VZ), d2vb = tx.get(kTabD2VZ), irc = tx.get(kTabIrc),
                    rcyl = tx.get(kTabRc);
  C<float, K> np_, dv, wx;
#pragma unroll
  for (int h = 0; h < K; ++h) {
    np_[h] = vis_nu * W.r[h];
    dv[h] = W.w[h] * irc[h];   // v_phi - v_bg
    wx[h] = W.dx0[h] * W.rad[h];
  }
  C<float, K> wm = prev1(wx);
  C<float, K> nm = prev1(np_), npp = next1(np_), tm = prev1(W.u), tp = next1(W.u), rm = prev1(W.v), rp = next1(W.v),
              vm = prev1(dv), vp = next1(dv);
#pragma unroll
  for (int h = 0; h < K; ++h) {
    const int c = cell_in<SL>(K, h);
    const float x2 = wm[h], x3 = wm[h] + wx[h];
    const float rho = W.rb[h] + W.r[h], n = vis_nu * rho;
    float dn, d1, d2;
    derivs<float>(0.0f, x2, x3, nm[h], np_[h], npp[h], dn, d2);
    dn += dnb[h];
    derivs<float>(0.0f, x2, x3, tm[h], W.u[h], tp[h], d1, d2);
    const float ddt = 2.0f * n * d2 + 2.0f * dn * d1;
    derivs<float>(0.0f, x2, x3, rm[h], W.v[h], rp[h], d1, d2);
    const float ddr = n * d2 + dn * d1;
    derivs<float>(0.0f, x2, x3, vm[h], dv[h], vp[h], d1, d2);
    const float ddv = n * (d2 + d2vb[h]) + dn * (d1 + d1vb[h]);
    if (c >= kNpad && c < lim - kNpad) {
      W.u[h] = W.u[h] + dt * ddt / rho;
      W.v[h] = W.v[h] + dt * ddr / rho;
      W.w[h] = W.w[h] + dt * rcyl[h] * ddv / rho;
    }
  }
}

}  // namespace pn
