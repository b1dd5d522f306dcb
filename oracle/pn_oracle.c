/*
 * pn_oracle.c -- FP64 CPU restatement of PEnGUIn's per-step hydro update.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is what results are compared with: tests/, the smoke()
 * entry and bench.py's cpu_baseline leg may load it; the product library
 * never links it and never falls back to it.
 *
 * What it restates (paths relative to the reference checkout):
 *   window structure of the sweeps ......... src/ppm/ppm.cu:26-307
 *   ppmlr stage order ...................... src/ppm/ppmlr.cu:2-56
 *   flatten ................................ src/ppm/flatten.cu:7-40
 *   paraset / parabola ..................... src/ppm/parabola.cu:70-176
 *   states ................................. src/ppm/states.cu:3-73
 *   riemann (EOS 0, 1, 2) .................. src/ppm/riemann.cu:3-115
 *   evolve ................................. src/ppm/evolve.cu:3-166
 *   remap / get_flux ....................... src/ppm/remap.cu:1-226
 *   forces / potential ..................... src/ppm/forces.cu:4-372
 *   viscosity (derivs, r, phi, theta) ...... src/ppm/viscosity.cu:2-19, 370-518
 *   Index_Shift ............................ src/orbital_advection/orbital_advection.cu:12-43
 *   rotate_rings ........................... src/ppm/ppm.cu:309-322
 *   CFL dt ................................. src/timestep/timestep.cu:12-149
 *   mass / planet-force reductions ......... src/output/device_func.cu:49-198
 *   grid / get_vol ......................... src/shared/grid.cu:67-248
 *   disk profile ........................... src/disk_profile/disk_profile.cu:9-171
 *   initial state and ghost rings .......... init.cpp:3-543
 *   planet ................................. src/planet/planet.cu:11-93
 *   step loop .............................. main.cu:270-387
 *
 * The reference runs one CUDA thread per cell of a 128-cell window with
 * __syncthreads() between stages.  Here each stage is a loop over the cells
 * of the window; per-thread registers become per-cell arrays and shared
 * arrays keep the reference's aliasing (the remap flux lives in the dx
 * scratch, the Riemann output reuses the sound-speed scratch), so every
 * value is computed from the same operands in the same order.  Built with
 * -ffp-contract=off and no fast-math.
 *
 * Deliberate choices where the reference is defective (each also noted where
 * it happens):
 *  - a pencil whose length is a multiple of 116 gets a full last window
 *    (the reference skips it, ppm.cu:34,48 / :169,196);
 *  - the CFL and mass reductions cover every ring (the reference stops at
 *    1024 rings, timestep.cu:99, device_func.cu:158);
 *  - ghost rings carry res_v = 0 and inc_j = 0 (never initialised in the
 *    reference, init.cpp:259-413).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "pn_oracle.h"

#define NPAD 6
#define ARR 128
#define REAL (ARR - 2 * NPAD)

static const double smallp = 1.0e-15;
static const double smallr = 1.0e-15;
static const double third = 0.3333333333333333333333333;
static const double fourthd = 1.3333333333333333333333333;
static const double twopi = 6.283185307179586476925286766559;
static const double pi_ = 3.1415926535897932384626433832795;
static const double EarthMass = 0.000003;

/* ------------------------------------------------------------------------ */
/* hydr_ring byte layout (variable_types.h:67-97) with a runtime jmax        */

typedef struct {
  double *r, *p, *u, *v, *w;
  double *x, *dx, *xvol, *xc;
  double *y, *dy, *yvol, *yc;
  double *z, *dz, *zvol, *zc;
  int32_t *i, *k;
  double *rot_v, *res_v;
  int32_t *rot_j, *inc_j;
} ring_t;

size_t orc_ring_bytes(int jmax) { return (size_t)72 * (size_t)jmax + 96; }

static ring_t ring_at(const void* base, int jmax, long idx) {
  char* b = (char*)base + (size_t)idx * orc_ring_bytes(jmax);
  size_t J = (size_t)jmax;
  ring_t R;
  R.r = (double*)(b);
  R.p = (double*)(b + 8 * J);
  R.u = (double*)(b + 16 * J);
  R.v = (double*)(b + 24 * J);
  R.w = (double*)(b + 32 * J);
  R.x = (double*)(b + 40 * J);
  R.dx = (double*)(b + 40 * J + 8);
  R.xvol = (double*)(b + 40 * J + 16);
  R.xc = (double*)(b + 40 * J + 24);
  R.y = (double*)(b + 40 * J + 32);
  R.dy = (double*)(b + 48 * J + 32);
  R.yvol = (double*)(b + 56 * J + 32);
  R.yc = (double*)(b + 64 * J + 32);
  R.z = (double*)(b + 72 * J + 32);
  R.dz = (double*)(b + 72 * J + 40);
  R.zvol = (double*)(b + 72 * J + 48);
  R.zc = (double*)(b + 72 * J + 56);
  R.i = (int32_t*)(b + 72 * J + 64);
  R.k = (int32_t*)(b + 72 * J + 68);
  R.rot_v = (double*)(b + 72 * J + 72);
  R.res_v = (double*)(b + 72 * J + 80);
  R.rot_j = (int32_t*)(b + 72 * J + 88);
  R.inc_j = (int32_t*)(b + 72 * J + 92);
  return R;
}

/* ------------------------------------------------------------------------ */
/* disk profile, disk_profile.cu:9-171                                       */

static double get_cs2(const pn_params* P, double r) { return P->sc_h * P->sc_h * pow(r, -P->p_beta); }
static double get_dcs2(const pn_params* P, double r) { return -P->p_beta * get_cs2(P, r) / r; }
static double get_h(const pn_params* P, double r) { return sqrt(get_cs2(P, r) * r * r * r); }
static double get_nu(const pn_params* P, double r) { return P->vis_nu * pow(r, P->p_alpha); }

static double get_rho1(const pn_params* P, double r) {
  double n;
  if (P->eos == 0) {
    n = -P->p_alpha;
  } else if (P->ndim == 3) {
    n = -P->p_alpha * (2.0 / (P->gam + 1.0));
  } else {
    n = -P->p_alpha;
  }
  return pow(r, n);
}

static double get_P1(const pn_params* P, double r) {
  if (P->eos == 0 || P->init_p_iso) return get_cs2(P, r) * get_rho1(P, r);
  return get_cs2(P, r) * pow(get_rho1(P, r), P->gam) / P->gam;
}

static double get_dP_dr1(const pn_params* P, double r) {
  double p1 = get_P1(P, r + 1.27245e-8);
  double p2 = get_P1(P, r - 1.27245e-8);
  return (p1 - p2) / 2.5449e-8;
}

static double get_verfac(const pn_params* P, double r, double z) {
  double cs2 = get_cs2(P, r);
  double rr = sqrt(r * r + z * z);
  return exp((1.0 / rr - 1.0 / r) / cs2);
}

static double get_rho2(const pn_params* P, double r, double z) {
  if (P->eos == 0) return get_rho1(P, r) * get_verfac(P, r, z);
  {
    double gamm = P->gam - 1.0;
    double gamfac2 = P->gam + 1.0;
    double h = sqrt(get_cs2(P, r) * r * r * r);
    double z_0 = sqrt(2.0 * h * h * pow(r, -P->p_alpha * (2.0 / gamfac2) * gamm) / gamm);
    if (z < z_0) {
      return pow(pow(get_rho1(P, r), gamm) - (gamm / (r * get_cs2(P, r))) * (1.0 - r / sqrt(r * r + z * z)),
                 1.0 / gamm);
    }
    return pow(10.0, log10(pow(pow(get_rho1(P, r), gamm) -
                                   (gamm / (r * get_cs2(P, r))) * (1.0 - r / sqrt(r * r + z_0 * z_0)),
                               1.0 / gamm)) -
                         6.0 * (z - z_0) / h);
  }
}

static double get_P2(const pn_params* P, double r, double z) {
  if (P->eos == 0 || P->init_p_iso) return get_cs2(P, r) * get_rho2(P, r, z);
  return get_cs2(P, r) * pow(get_rho2(P, r, z), P->gam) / P->gam;
}

static double get_dP_dr2(const pn_params* P, double r, double z) {
  double p1 = get_P2(P, r + 1.27e-8, z);
  double p2 = get_P2(P, r - 1.27e-8, z);
  return (p1 - p2) / 2.54e-8;
}

static double get_dP_s(const pn_params* P, double r, double pol) {
  double dr = 1.27e-7;
  double sinpol = sin(pol);
  double cospol = cos(pol);
  double r1 = (r + dr) * sinpol;
  double z1 = (r + dr) * cospol;
  double r2 = (r - dr) * sinpol;
  double z2 = (r - dr) * cospol;
  return (get_P2(P, r1, z1) - get_P2(P, r2, z2)) / (2.0 * dr);
}

static double get_viscous_vr(const pn_params* P, double r) {
  if (P->visc_flag == 1) return -1.5 * get_nu(P, r) / r;
  return 0.0;
}

/* unused in the live reference but part of its disk module */
double orc_unused_dcs2(const pn_params* P, double r) { return get_dcs2(P, r); }

/* ------------------------------------------------------------------------ */
/* grid.cu:67-248                                                            */

void orc_grid(int nzones, double xmin, double xmax, int lg, double* xa, double* xc, double* dx) {
  if (lg == 0) {
    double dxfac = (xmax - xmin) / (double)nzones;
    for (int n = 0; n < nzones; n++) {
      xa[n] = xmin + ((double)n) * dxfac;
      dx[n] = dxfac;
      xc[n] = xa[n] + 0.5 * dx[n];
    }
  } else {
    /* lg 1, 2, 3 are the same logarithmic spacing in the reference */
    double lg_dxfac = (log(xmax) - log(xmin)) / (double)nzones;
    xa[0] = xmin;
    for (int n = 1; n < nzones; n++) {
      xa[n] = xmin * exp(lg_dxfac * ((double)n));
      dx[n - 1] = xa[n] - xa[n - 1];
      xc[n - 1] = xa[n - 1] + 0.5 * dx[n - 1];
    }
    dx[nzones - 1] = xmax - xa[nzones - 1];
    xc[nzones - 1] = xa[nzones - 1] + 0.5 * dx[nzones - 1];
  }
}

static double get_vol(const pn_params* P, double xa, double dx, int axis) {
  double vol = 0.0;
  if (axis == 0) {
    if (P->ngeomx == 0) vol = dx;
    else if (P->ngeomx == 1) vol = dx * (xa + 0.5 * dx);
    else if (P->ngeomx == 2) vol = dx * (xa * (xa + dx) + dx * dx * third);
  } else if (axis == 1) {
    vol = dx; /* ngeomy 0, 3, 4 */
  } else {
    if (P->ndim == 2) vol = 1.0;
    else if (P->ngeomz == 0) vol = dx;
    else if (P->ngeomz == 5) vol = cos(xa) - cos(xa + dx);
  }
  return vol;
}

/* ------------------------------------------------------------------------ */
/* seeded noise                                                              */

uint64_t orc_splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double orc_noise_u(uint64_t seed, uint64_t cell) {
  uint64_t b = orc_splitmix64(seed ^ cell);
  return (double)(b >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

/* ------------------------------------------------------------------------ */
/* initial state, init.cpp:3-543                                             */

static void fill_phi(const pn_params* P, ring_t R, int jmax, const double* ya, const double* yc,
                     const double* dy, int from_edge) {
  for (int j = 0; j < jmax; j++) {
    R.y[j] = ya[j];
    R.dy[j] = dy[j];
    /* init_cells_val recomputes yc from y and dy (init.cpp:205); the ghost
       rings copy zyc (init.cpp:318) -- the same value */
    R.yc[j] = from_edge ? R.y[j] + 0.5 * R.dy[j] : yc[j];
    R.yvol[j] = get_vol(P, ya[j], dy[j], 1);
  }
}

int orc_init_disk(const pn_grid* g, const pn_params* P, uint64_t seed, double noise, void* rings,
                  void* lft, void* rgh, void* udr, void* top, pn_symdisk* val) {
  int imax = g->imax, jmax = g->jmax, kmax = g->kmax;
  if (!g || !P || !rings || !lft || !rgh || !val || imax < 1 || jmax < 2 * NPAD || kmax < 1) return PN_ERR_ARG;
  if (P->ndim == 3 && (!udr || !top)) return PN_ERR_ARG;
  double *zxa = calloc(imax, 8), *zxc = calloc(imax, 8), *zdx = calloc(imax, 8);
  double *zya = calloc(jmax, 8), *zyc = calloc(jmax, 8), *zdy = calloc(jmax, 8);
  double *zza = calloc(kmax, 8), *zzc = calloc(kmax, 8), *zdz = calloc(kmax, 8);
  orc_grid(imax, P->xmin, P->xmax, P->grspx, zxa, zxc, zdx);
  orc_grid(jmax, P->ymin, P->ymax, P->grspy, zya, zyc, zdy);
  orc_grid(kmax, P->zmin, P->zmax, P->grspz, zza, zzc, zdz);

  /* init_den, init.cpp:92-138 */
  for (int i = 0; i < imax; i++)
    for (int k = 0; k < kmax; k++) {
      double r = 0.0, z = 0.0;
      if (P->ngeomz == 5) {
        r = zxc[i] * sin(zzc[k]);
        z = zxc[i] * cos(zzc[k]);
      } else {
        r = zxc[i];
        z = zzc[k];
      }
      if (P->ndim == 3) {
        val[i + imax * k].r = get_rho2(P, r, z);
        val[i + imax * k].p = get_P2(P, r, z);
      } else {
        val[i + imax * k].r = get_rho1(P, r);
        val[i + imax * k].p = get_P1(P, r);
      }
    }
  /* init_speed, init.cpp:142-183 */
  for (int i = 0; i < imax; i++)
    for (int k = 0; k < kmax; k++) {
      double r = 0.0, z = 0.0, vk2 = 0.0, dP = 0.0;
      if (P->ndim == 2) {
        r = zxc[i];
        z = 0.0;
        vk2 = 1.0 / r;
        dP = get_dP_dr1(P, zxc[i]);
      } else if (P->ngeomz == 5) {
        vk2 = 1.0 / zxc[i];
        r = zxc[i] * sin(zzc[k]);
        z = zxc[i] * cos(zzc[k]);
        dP = get_dP_s(P, zxc[i], zzc[k]);
      } else {
        r = zxc[i];
        z = zzc[k];
        dP = get_dP_dr2(P, zxc[i], zzc[k]);
        vk2 = r * r * pow(r * r + z * z, -1.5);
      }
      (void)z;
      vk2 += zxc[i] * dP / val[i + imax * k].r;
      val[i + imax * k].v = sqrt(vk2);
      val[i + imax * k].u = get_viscous_vr(P, r);
      val[i + imax * k].w = 0.0;
    }

  /* init_cells, init.cpp:187-249, plus the seeded noise */
  size_t rb = orc_ring_bytes(jmax);
  memset(rings, 0, rb * (size_t)imax * (size_t)kmax);
  for (int k = 0; k < kmax; k++)
    for (int i = 0; i < imax; i++) {
      long m = i + (long)imax * k;
      ring_t R = ring_at(rings, jmax, m);
      int idx = i + imax * k;
      *R.k = k;
      *R.z = zza[k];
      *R.dz = zdz[k];
      *R.zc = *R.z + 0.5 * *R.dz;
      *R.i = i;
      *R.x = zxa[i];
      *R.dx = zdx[i];
      *R.xc = *R.x + 0.5 * *R.dx;
      double rad_cyl = *R.xc;
      if (P->ngeomz == 5) rad_cyl *= sin(*R.zc);
      fill_phi(P, R, jmax, zya, zyc, zdy, 1);
      for (int j = 0; j < jmax; j++) {
        double rho = val[idx].r;
        if (noise != 0.0) rho = rho * (1.0 + noise * orc_noise_u(seed, (uint64_t)j + (uint64_t)jmax * (uint64_t)idx));
        R.r[j] = rho;
        if (noise != 0.0) {
          /* re-derive the pressure of the perturbed surface density with the
             closure the sweeps store (ppm.cu:139) */
          if (P->eos == 0 || P->init_p_iso) R.p[j] = rho * get_cs2(P, rad_cyl);
          else R.p[j] = get_cs2(P, rad_cyl) * pow(rho, P->gam) / P->gam;
        } else {
          R.p[j] = val[idx].p;
        }
        R.u[j] = val[idx].u;
        R.v[j] = val[idx].v;
        R.w[j] = 0.0;
      }
      *R.rot_j = 0;
      *R.inc_j = 0;
      *R.rot_v = 0.0; /* pinned host allocation, assumed zero (main.cu:192) */
      *R.res_v = 0.0;
      *R.xvol = get_vol(P, *R.x, *R.dx, 0);
      *R.zvol = get_vol(P, *R.z, *R.dz, 2);
    }

  /* init_bound, init.cpp:253-413 (fixed radial ghost rings) */
  memset(lft, 0, rb * NPAD * (size_t)kmax);
  memset(rgh, 0, rb * NPAD * (size_t)kmax);
  const double FrRot_init = 1.0; /* init.cpp:323,404,467,530 hard-code a unit frame rate */
  for (int side = 0; side < 2; side++) {
    for (int i = 0; i < NPAD; i++)
      for (int k = 0; k < kmax; k++) {
        int ib = i + NPAD * k;
        ring_t R = ring_at(side == 0 ? lft : rgh, jmax, ib);
        if (side == 0) {
          *R.x = zxa[0] - zdx[0] * (double)(NPAD - i);
          *R.dx = zdx[0];
        } else {
          *R.x = zxa[imax - 1] + zdx[imax - 1] * (double)(i + 1);
          *R.dx = zdx[imax - 1];
        }
        *R.xc = *R.x + 0.5 * *R.dx;
        *R.xvol = get_vol(P, *R.x, *R.dx, 0);
        *R.z = zza[k];
        *R.dz = zdz[k];
        *R.zc = zzc[k];
        *R.zvol = get_vol(P, zza[k], zdz[k], 2);
        *R.i = 0;
        *R.k = 0;
        double rr = *R.xc, r = rr, z = 0.0, rho = 0.0, Pr = 0.0, dP = 0.0, vk = 0.0;
        if (P->ndim == 2) {
          r = rr;
          z = 0.0;
          rho = get_rho1(P, r);
          Pr = get_P1(P, r);
          dP = get_dP_dr1(P, r);
          vk = 1.0 / r;
        } else if (P->ngeomz == 5) {
          r = rr * sin(zzc[k]);
          z = rr * cos(zzc[k]);
          /* the reference's left ghost ring leaves rho and P from the
             previous assignment (init.cpp:288-294); the right one sets them
             (init.cpp:365-371).  Both sides use the 3-D profile here. */
          rho = get_rho2(P, r, z);
          Pr = get_P2(P, r, z);
          dP = get_dP_s(P, rr, zzc[k]);
          vk = 1.0 / rr;
        } else {
          r = rr;
          z = zzc[k];
          rho = get_rho2(P, r, z);
          Pr = get_P2(P, r, z);
          dP = get_dP_dr2(P, r, z);
          vk = r * r * pow(r * r + z * z, -1.5);
        }
        vk += rr * dP / rho;
        vk = sqrt(vk);
        for (int j = 0; j < jmax; j++) {
          R.r[j] = rho;
          R.p[j] = Pr;
          R.u[j] = get_viscous_vr(P, r);
          R.v[j] = vk;
          R.w[j] = 0.0;
        }
        fill_phi(P, R, jmax, zya, zyc, zdy, 0);
        if (P->fargo_flag == 1) *R.rot_v = vk - FrRot_init * *R.xc;
        else *R.rot_v = 0.0;
        *R.rot_j = 0;
        *R.res_v = 0.0;
        *R.inc_j = 0;
      }
  }

  /* theta ghost rings (init.cpp:415-541), 3-D only */
  if (P->ndim == 3) {
    memset(udr, 0, rb * NPAD * (size_t)imax);
    memset(top, 0, rb * NPAD * (size_t)imax);
    for (int side = 0; side < 2; side++)
      for (int k = 0; k < NPAD; k++)
        for (int i = 0; i < imax; i++) {
          int ib = k + NPAD * i;
          ring_t R = ring_at(side == 0 ? udr : top, jmax, ib);
          if (side == 0) {
            *R.z = zza[0] - zdz[0] * (double)(NPAD - k);
            *R.dz = zdz[0];
          } else {
            *R.z = zza[kmax - 1] + zdz[kmax - 1] * (double)(k + 1);
            *R.dz = zdz[kmax - 1];
          }
          *R.zc = *R.z + 0.5 * *R.dz;
          *R.zvol = get_vol(P, *R.z, *R.dz, 2);
          *R.x = zxa[i];
          *R.dx = zdx[i];
          *R.xc = zxc[i];
          *R.xvol = get_vol(P, zxa[i], zdx[i], 0);
          double rr = zxc[i], r, z, dP;
          if (P->ngeomz == 5) {
            z = rr * cos(*R.zc);
            r = sqrt(rr * rr - z * z);
            dP = get_dP_s(P, rr, *R.zc);
          } else {
            r = rr;
            z = *R.zc;
            dP = get_dP_dr2(P, r, z);
          }
          double rho = get_rho2(P, r, z);
          double Pr = get_P2(P, r, z);
          double vk = 1.0 / rr;
          vk += rr * dP / rho;
          vk = sqrt(vk);
          for (int j = 0; j < jmax; j++) {
            R.r[j] = rho;
            R.p[j] = Pr;
            R.u[j] = get_viscous_vr(P, r);
            R.v[j] = vk;
            R.w[j] = 0.0;
          }
          fill_phi(P, R, jmax, zya, zyc, zdy, 0);
          if (P->fargo_flag == 1) *R.rot_v = vk - FrRot_init * *R.xc;
          else *R.rot_v = 0.0;
          *R.rot_j = 0;
          *R.res_v = 0.0;
          *R.inc_j = 0;
          *R.i = 0;
          *R.k = 0;
        }
  }
  free(zxa); free(zxc); free(zdx); free(zya); free(zyc); free(zdy); free(zza); free(zzc); free(zdz);
  return PN_OK;
}

/* ------------------------------------------------------------------------ */
/* forces.cu:4-372 (bary_flag undefined)                                     */

double orc_potential(const pn_params* P, const pn_body* p, double rad, double azi, double z) {
  if (P->plnt_flag > 0) {
    double cosfac = cos(azi - p->y);
    double stm = 1.0 / (1.0 + p->m);
    double plm = p->m * stm;
    double plx = p->x * stm;
    double stx = p->x * plm;
    double Rp = rad * rad + plx * plx - 2.0 * plx * rad * cosfac + z * z;
    Rp = sqrt(Rp + p->rs * p->rs);
    double Rs = sqrt(rad * rad + stx * stx + 2.0 * stx * rad * cosfac + z * z);
    return -plm / Rp - stm / Rs;
  }
  return -1.0 / sqrt(rad * rad + z * z);
}

/* star_planet_grav_rad / _azi / _pol (forces.cu:23-110), spherical.  Without bary_flag
   the star sits at the origin (its pull -stm/r^2 and the indirect term of the planet
   included); with it star and planet orbit their barycentre (star at -x m/(1+m), planet at
   x/(1+m)) and there is no indirect term. */
static double grav_rad(const pn_body* p, int bary, double rad, double azi, double pol, double dt) {
  double cosfac = cos(azi - (p->y + dt * p->vy));
  double sinpol = sin(pol);
  double stm = 1.0 / (1.0 + p->m);
  double plm = p->m * stm;
  if (bary) {
    double plx = p->x * stm, stx = p->x * plm;
    double Rp = rad * rad + plx * plx - 2.0 * plx * rad * cosfac * sinpol;
    Rp = sqrt(Rp + p->rs * p->rs);
    double Rs = sqrt(rad * rad + stx * stx + 2.0 * stx * rad * cosfac * sinpol);
    return -plm * (rad - plx * cosfac * sinpol) / (Rp * Rp * Rp) - stm * (rad + stx * cosfac * sinpol) / (Rs * Rs * Rs);
  }
  double plx = p->x;
  double Rp = rad * rad + plx * plx - 2.0 * plx * rad * cosfac * sinpol;
  Rp = sqrt(Rp + p->rs * p->rs);
  return -plm * (rad - plx * cosfac * sinpol) / (Rp * Rp * Rp) - stm / (rad * rad) -
         plm * sinpol * cosfac / (plx * plx);
}

static double grav_azi(const pn_body* p, int bary, double rad, double azi, double pol, double dt) {
  double cosfac = cos(azi - (p->y + dt * p->vy));
  double sinfac = sin(azi - (p->y + dt * p->vy));
  double sinpol = sin(pol);
  rad /= sinpol;
  double stm = 1.0 / (1.0 + p->m);
  double plm = p->m * stm;
  if (bary) {
    double plx = p->x * stm, stx = p->x * plm;
    double Rp = rad * rad + plx * plx - 2.0 * plx * rad * cosfac * sinpol;
    Rp = sqrt(Rp + p->rs * p->rs);
    double Rs = sqrt(rad * rad + stx * stx + 2.0 * stx * rad * cosfac * sinpol);
    return -plm * plx * sinfac / (Rp * Rp * Rp) + stm * stx * sinfac / (Rs * Rs * Rs);
  }
  double plx = p->x;
  double Rp = rad * rad + plx * plx - 2.0 * plx * rad * cosfac * sinpol;
  Rp = sqrt(Rp + p->rs * p->rs);
  return -plm * plx * sinfac / (Rp * Rp * Rp) + plm * sinpol * sinfac / plx / plx;
}

static double grav_pol(const pn_body* p, int bary, double rad, double azi, double pol, double dt) {
  double cosfac = cos(azi - (p->y + dt * p->vy));
  double cospol = cos(pol);
  double sinpol = sin(pol);
  double stm = 1.0 / (1.0 + p->m);
  double plm = p->m * stm;
  if (bary) {
    double plx = p->x * stm, stx = p->x * plm;
    double Rp = rad * rad + plx * plx - 2.0 * plx * rad * cosfac * sinpol;
    Rp = sqrt(Rp + p->rs * p->rs);
    double Rs = sqrt(rad * rad + stx * stx + 2.0 * stx * rad * cosfac * sinpol);
    return plm * plx * cosfac * cospol / (Rp * Rp * Rp) - stm * stx * cosfac * cospol / (Rs * Rs * Rs);
  }
  double plx = p->x;
  double Rp = rad * rad + plx * plx - 2.0 * plx * rad * cosfac * sinpol;
  Rp = sqrt(Rp + p->rs * p->rs);
  return plm * plx * cosfac * cospol / (Rp * Rp * Rp) - plm * cosfac * cospol / (sinpol * plx * plx);
}

/* star_planet_grav_*_cyl (forces.cu:112-182), cylindrical; without bary_flag the planet with
   its indirect term (the star's -1/r^2 is added by the caller), with it star and planet
   about their barycentre. */
static double grav_rad_cyl(const pn_body* p, int bary, double rad, double azi, double z, double dt) {
  double cosfac = cos(azi - (p->y + dt * p->vy));
  if (bary) {
    double stm = 1.0 / (1.0 + p->m), plm = p->m * stm, plx = p->x * stm, stx = p->x * plm;
    double Rp = sqrt(rad * rad + plx * plx - 2.0 * plx * rad * cosfac + z * z + p->rs * p->rs);
    double Rs = sqrt(rad * rad + stx * stx + 2.0 * stx * rad * cosfac + z * z);
    return -plm * (rad - plx * cosfac) / (Rp * Rp * Rp) - stm * (rad + stx * cosfac) / (Rs * Rs * Rs);
  }
  double plm = p->m;
  double plx = p->x;
  double Rp = sqrt(rad * rad + plx * plx - 2.0 * plx * rad * cosfac + z * z + p->rs * p->rs);
  return -plm * (rad - plx * cosfac) / (Rp * Rp * Rp) - plm * cosfac / (plx * plx);
}

static double grav_azi_cyl(const pn_body* p, int bary, double rad, double azi, double z, double dt) {
  double cosfac = cos(azi - (p->y + dt * p->vy));
  double sinfac = sin(azi - (p->y + dt * p->vy));
  if (bary) {
    double stm = 1.0 / (1.0 + p->m), plm = p->m * stm, plx = p->x * stm, stx = p->x * plm;
    double Rp = sqrt(rad * rad + plx * plx - 2.0 * plx * rad * cosfac + z * z + p->rs * p->rs);
    double Rs = sqrt(rad * rad + stx * stx + 2.0 * stx * rad * cosfac + z * z);
    return -plm * plx * sinfac / (Rp * Rp * Rp) + stm * stx * sinfac / (Rs * Rs * Rs);
  }
  double plm = p->m;
  double plx = p->x;
  double Rp = sqrt(rad * rad + plx * plx - 2.0 * plx * rad * cosfac + z * z + p->rs * p->rs);
  return -plm * plx * sinfac / (Rp * Rp * Rp) + plm * sinfac / plx / plx;
}

static double grav_pol_cyl(const pn_body* p, int bary, double rad, double azi, double z, double dt) {
  double cosfac = cos(azi - (p->y + dt * p->vy));
  if (bary) {
    double stm = 1.0 / (1.0 + p->m), plm = p->m * stm, plx = p->x * stm, stx = p->x * plm;
    double Rp = sqrt(rad * rad + plx * plx - 2.0 * plx * rad * cosfac + z * z + p->rs * p->rs);
    double Rs = sqrt(rad * rad + stx * stx + 2.0 * stx * rad * cosfac + z * z);
    return -plm * z / (Rp * Rp * Rp) - stm * z / (Rs * Rs * Rs);
  }
  double plm = p->m;
  double plx = p->x;
  double Rp = sqrt(rad * rad + plx * plx - 2.0 * plx * rad * cosfac + z * z + p->rs * p->rs);
  return -plm * z / (Rp * Rp * Rp);
}

double orc_force_rad_cyl(const pn_body* p, double rad, double azi, double z, double dt) {
  return grav_rad_cyl(p, 0, rad, azi, z, dt);
}
double orc_force_azi_cyl(const pn_body* p, double rad, double azi, double z, double dt) {
  return grav_azi_cyl(p, 0, rad, azi, z, dt);
}
/* the gravity the sweeps apply, component comp (0 r, 1 phi, 2 z / theta) in geometry geo
   (0 cylindrical, 1 spherical), with or without the barycentric frame; cylindrical r
   includes the star's -1/r^2 outside the barycentric frame, as get_fx adds it */
double orc_gravity(int geo, int comp, int bary, const pn_body* p, double rad, double azi, double pol, double dt) {
  if (geo == 0) {
    if (comp == 0) return grav_rad_cyl(p, bary, rad, azi, pol, dt) + (bary ? 0.0 : -1.0 / rad / rad);
    if (comp == 1) return grav_azi_cyl(p, bary, rad, azi, pol, dt);
    return grav_pol_cyl(p, bary, rad, azi, pol, dt);
  }
  if (comp == 0) return grav_rad(p, bary, rad, azi, pol, dt);
  if (comp == 1) return grav_azi(p, bary, rad, azi, pol, dt);
  return grav_pol(p, bary, rad, azi, pol, dt);
}

/* get_fx_bound / get_fx (forces.cu:185-263): interface (bound) forms use the
   mean of the two neighbouring transverse momenta */
static void get_fx(const pn_params* P, int bound, int n, double dt, double rad, double azi, double pol,
                   double* grav, double* ftot, const pn_body* pl, const double* v, const double* w) {
  if (P->ngeomx == 0) {
    *grav = 0.0;
    *ftot = *grav;
  } else if (P->ngeomx == 1) {
    if (P->plnt_flag > 0) {
      *grav = grav_rad_cyl(pl, P->bary_flag, rad, azi, pol, dt);
      if (!P->bary_flag) *grav += -1.0 / rad / rad;
    } else {
      double dis = sqrt(rad * rad + pol * pol);
      *grav = -rad / dis / dis / dis;
    }
    *ftot = *grav;
    if (bound) *ftot += pow(0.5 * (v[n] + v[n - 1]), 2) / rad / rad / rad;
    else *ftot += v[n] * v[n] / rad / rad / rad;
  } else if (P->ngeomx == 2) {
    if (P->plnt_flag > 0) *grav = grav_rad(pl, P->bary_flag, rad, azi, pol, dt);
    else *grav = -1.0 / rad / rad;
    *ftot = *grav;
    if (bound) {
      *ftot += pow(0.5 * (w[n] + w[n - 1]), 2) / rad / rad / rad;
      *ftot += pow(0.5 * (v[n] + v[n - 1]), 2) / rad / pow(rad * sin(pol), 2);
    } else {
      *ftot += w[n] * w[n] / rad / rad / rad;
      *ftot += v[n] * v[n] / rad / pow(rad * sin(pol), 2);
    }
  }
}

static void get_fy(const pn_params* P, double dt, double rad, double azi, double pol, double* grav,
                   double* ftot, const pn_body* pl) {
  if (P->ngeomy == 3) {
    *grav = (P->plnt_flag > 0) ? grav_azi_cyl(pl, P->bary_flag, rad, azi, pol, dt) : 0.0;
  } else if (P->ngeomy == 4) {
    *grav = (P->plnt_flag > 0) ? grav_azi(pl, P->bary_flag, rad, azi, pol, dt) : 0.0;
  } else {
    *grav = 0.0;
  }
  *ftot = *grav;
}

static void get_fz(const pn_params* P, int bound, int n, double dt, double rad, double azi, double pol,
                   double* grav, double* ftot, const pn_body* pl, const double* w) {
  if (P->ngeomz == 0) {
    if (P->plnt_flag > 0) {
      *grav = grav_pol_cyl(pl, P->bary_flag, rad, azi, pol, dt);
    } else {
      double dis = sqrt(rad * rad + pol * pol);
      *grav = -pol / dis / dis / dis;
    }
    *ftot = *grav;
  } else if (P->ngeomz == 5) {
    *grav = (P->plnt_flag > 0) ? grav_pol(pl, P->bary_flag, rad, azi, pol, dt) : 0.0;
    double rad_cyl = rad * sin(pol);
    *ftot = *grav;
    if (bound) *ftot += (pow(0.5 * (w[n] + w[n - 1]), 2) * cos(pol)) / rad_cyl / rad_cyl / rad_cyl;
    else *ftot += (w[n] * w[n] * cos(pol)) / rad_cyl / rad_cyl / rad_cyl;
  } else {
    *grav = 0.0;
    *ftot = 0.0;
  }
}

/* ------------------------------------------------------------------------ */
/* window scratch: the reference's shared arrays and per-thread registers   */

typedef struct {
  /* shared (ppmlr.cu:7-11) */
  double q[ARR], dx[ARR], umid[ARR], pmid[ARR], tmp1[ARR];
  /* paraset outputs per thread */
  double c0[ARR], c1[ARR], c2[ARR], c3[ARR], c4[ARR];
  /* paraset registers carried from stage 1 to stage 2 */
  double dxn_r[ARR], ain_r[ARR], bi_r[ARR], c_r[ARR], ci_r[ARR];
  /* parabola registers */
  double pa[ARR], pdar[ARR], pt1[ARR], par_[ARR];
  /* states */
  double flat[ARR], ftot_s[ARR];
  double Spl[ARR], Srl[ARR], Sul[ARR], Spr[ARR], Srr[ARR], Sur[ARR];
  /* evolve */
  double delx[ARR], dtdm[ARR], xa[ARR], dvol[ARR], fto1[ARR], fto2[ARR], gra1[ARR], gra2[ARR], xcl[ARR];
  /* remap */
  double lvol[ARR], dxf[ARR], dxf2[ARR], fr[ARR], dxn[ARR], dxnm[ARR], m[ARR], invm[ARR];
  int delC[ARR];
  /* viscosity */
  double nr[ARR], ddu[ARR], ddv[ARR], ddw[ARR], rho[ARR];
} win_t;

/* min3 / max3 (simple_func.cu:2-28) */
static double min3(double A, double B, double C) {
  if (A < B) return (C < A) ? C : A;
  return (C < B) ? C : B;
}
static double max3(double A, double B, double C) {
  if (A > B) return (C > A) ? C : A;
  return (C > B) ? C : B;
}

/* paraset, parabola.cu:138-176 */
static void paraset(win_t* W, int nmin, int nmax, const double* dx, double* a, double* ai, double* b) {
  for (int n = nmin; n < nmax - 1; n++) {
    double dx_n = dx[n];
    double an = dx_n + dx[n + 1];
    double ain = 1.0 / an;
    double bn = an + dx_n;
    double bi = 1.0 / bn;
    double c = an + dx[n + 1];
    double ci = 1.0 / c;
    a[n] = an;
    ai[n] = ain;
    b[n] = bn;
    W->dxn_r[n] = dx_n;
    W->ain_r[n] = ain;
    W->bi_r[n] = bi;
    W->c_r[n] = c;
    W->ci_r[n] = ci;
  }
  for (int n = nmin + 1; n < nmax - 2; n++) {
    double d = 1.0 / (a[n - 1] + a[n + 1]);
    double dx_n = W->dxn_r[n], ain = W->ain_r[n], bi = W->bi_r[n], ci = W->ci_r[n];
    W->c0[n] = dx_n * ain * (1.0 + 2.0 * d * dx[n + 1] * (a[n - 1] * bi - a[n + 1] * ci));
    W->c1[n] = -d * dx_n * a[n - 1] * bi;
    W->c2[n] = d * dx[n + 1] * a[n + 1] * ci;
  }
  for (int n = nmin + 1; n < nmax - 1; n++) {
    double dx_n = W->dxn_r[n], ain = W->ain_r[n], c = W->c_r[n];
    double d = dx_n / (a[n - 1] + dx[n + 1]);
    W->c3[n] = d * b[n - 1] * ain;
    W->c4[n] = d * c * ai[n - 1];
  }
}

/* parabola, parabola.cu:70-136 */
static void parabola(win_t* W, int nmin, int nmax, const double* flat, const double* val, double* al,
                     double* da, double* a6) {
  for (int n = nmin + 1; n < nmax - 1; n++) {
    double a = val[n];
    double da_r = val[n + 1] - a;
    double da_l = a - val[n - 1];
    double t1 = W->c3[n] * da_r + W->c4[n] * da_l;
    if (t1 < 0.0) t1 = -min3(-t1, 2.0 * fabs(da_l), 2.0 * fabs(da_r));
    else t1 = min3(t1, 2.0 * fabs(da_l), 2.0 * fabs(da_r));
    if (da_l * da_r < 0.0) t1 = 0.0;
    a6[n] = t1;
    W->pa[n] = a;
    W->pdar[n] = da_r;
    W->pt1[n] = t1;
  }
  for (int n = nmin + 1; n < nmax - 2; n++) {
    double ar = W->pa[n] + W->c0[n] * W->pdar[n] + W->c1[n] * a6[n + 1] + W->c2[n] * W->pt1[n];
    W->par_[n] = ar;
  }
  /* the writes al[n+1] = ar happen after every thread has read a6[n+1]
     (reads precede the barrier): apply them as a second pass */
  for (int n = nmin + 1; n < nmax - 2; n++) al[n + 1] = W->par_[n];
  for (int n = nmin + 2; n < nmax - 2; n++) {
    double a = W->pa[n];
    double t1 = flat[n] * a;
    double t2 = 1.0 - flat[n];
    double ar = t1 + t2 * W->par_[n];
    double la = t1 + t2 * al[n];
    t1 = ar - la;
    t2 = t1 * t1;
    t1 = t1 * (6.0 * (a - 0.5 * (ar + la)));
    if ((ar - a) * (a - la) <= 0.0) {
      ar = a;
      la = a;
    }
    if (t2 < t1) la = 3.0 * a - 2.0 * ar;
    if (t2 < -t1) ar = 3.0 * a - 2.0 * la;
    W->par_[n] = ar; /* scratch, no longer needed as the stage-2 value */
    W->pt1[n] = la;
    da[n] = ar - la;
    a6[n] = 6.0 * (a - 0.5 * (ar + la));
  }
  for (int n = nmin + 2; n < nmax - 2; n++) al[n] = W->pt1[n];
}

/* flatten, flatten.cu:7-40 */
static void flatten(win_t* W, int nmax, const double* p, const double* u, double* steep) {
  for (int n = 0; n < nmax; n++) {
    W->flat[n] = 0.0;
    steep[n] = 0.0;
  }
  for (int n = 2; n < nmax - 2; n++) {
    int shock = 0;
    double dp1 = p[n + 1] - p[n - 1];
    double dp2 = p[n + 2] - p[n - 2];
    if (fabs(dp2) < smallp) dp2 = smallp;
    double ddp = fabs(dp1) / fmin(p[n + 1], p[n - 1]);
    if (ddp > 0.33) shock = 1;
    if (u[n - 1] < u[n + 1]) shock = 0;
    if (shock) {
      ddp = (dp1 / dp2 - 0.75) * 5.0;
      steep[n] = fmax(0.0, ddp);
    }
  }
  for (int n = 2; n < nmax - 2; n++) {
    double dp1 = max3(steep[n - 1], steep[n], steep[n + 1]);
    W->flat[n] = fmin(0.5, dp1);
  }
}

/* the energy a cell carries into the Lagrangian step: EOS 2 as ppm.cu:106-108 forms it (v is
   the first transverse slot, already r_cyl v_phi in the radial sweep); EOS 3 the internal
   energy plus the sweep direction's kinetic energy; unused (0) otherwise */
static double orc_energy(const pn_params* P, double p, double r, double u, double v) {
  if (P->eos == 2) return p / (r * (P->gam - 1.0)) + 0.5 * ((u * u) + (v * v));
  if (P->eos == 3) return p / (r * (P->gam - 1.0)) + 0.5 * (u * u);
  return 0.0;
}

/* riemann, riemann.cu:3-115 */
void orc_riemann(int eos, double gam, double pl, double rl, double ul, double pr, double rr, double ur,
                 double* umid, double* pmid) {
  double gamm = gam - 1.0, gamfac2 = gam + 1.0, gamfac1 = gamfac2 / gam / 2.0, gamz = gamm / gam / 2.0;
  if (eos == 0) {
    double t1 = 0.5 * (sqrt(pl / rl) + sqrt(pr / rr));
    double t2 = -0.5 * (ur - ul) / t1;
    double pm = sqrt(pl * pr) * exp(t2);
    *umid = ul - t1 * log(pm / pl);
    *pmid = fmax(smallp, pm);
  } else if (eos == 2 || eos == 3) {
    double cs_l = sqrt(gam * pl / rl);
    double cs_r = sqrt(gam * pr / rr);
    double pm = pow((cs_l + cs_r - (gamm / 2.0) * (ur - ul)) / (cs_l * pow(pl, -gamz) + cs_r * pow(pr, -gamz)),
                    1.0 / gamz);
    double t1, t2;
    if (pm > pr) t1 = (pm - pr) * sqrt(2.0 / (rr * (gamfac2 * pm + pr * gamm)));
    else t1 = 2.0 * cs_r * (pow(pm / pr, gamz) - 1.0) / gamm;
    if (pm > pl) t2 = (pm - pl) * sqrt(2.0 / (rl * (gamfac2 * pm + pl * gamm)));
    else t2 = 2.0 * cs_l * (pow(pm / pl, gamz) - 1.0) / gamm;
    *umid = 0.5 * (ur + ul) + 0.5 * (t1 - t2);
    *pmid = fmax(smallp, pm);
  } else {
    double cs_l = sqrt(gam * pl * rl);
    double cs_r = sqrt(gam * pr * rr);
    rl = 1.0 / rl;
    rr = 1.0 / rr;
    double pm = pr - pl - cs_r * (ur - ul);
    pm = pl + pm * cs_l / (cs_l + cs_r);
    pm = fmax(smallp, pm);
    double t1 = 0.0, t2 = 0.0, en_l, en_r, uml, umr, dum;
    for (int l = 0; l < 10; l++) {
      dum = pm;
      t1 = 1.0 + gamfac1 * (pm - pl) / pl;
      t2 = 1.0 + gamfac1 * (pm - pr) / pr;
      t1 = cs_l * sqrt(t1);
      t2 = cs_r * sqrt(t2);
      en_l = 4.0 * rl * t1 * t1;
      en_r = 4.0 * rr * t2 * t2;
      en_l = -en_l * t1 / (en_l - gamfac2 * (pm - pl));
      en_r = en_r * t2 / (en_r - gamfac2 * (pm - pr));
      uml = ul - (pm - pl) / t1;
      umr = ur + (pm - pr) / t2;
      pm = pm + (umr - uml) * (en_r * en_l) / (en_r - en_l);
      pm = fmax(smallp, pm);
      if (fabs(pm - dum) / pm < 1.0e-8) l = 10;
    }
    uml = ul - (pm - pl) / t1;
    umr = ur + (pm - pr) / t2;
    *umid = 0.5 * (uml + umr);
    *pmid = fmax(smallp, pm);
  }
}

/* ppmlr on one window: ppmlr.cu:2-56 with states, riemann, evolve, remap */
typedef struct {
  const pn_params* P;
  int axis, nmax;
  double dt, FrRot;
  const pn_body* pl;
  const double *rad, *azi, *pol, *v_rot, *dvol0;
  double *xa0, *dx0;
  double *r, *p, *u, *v, *w, *e;
} wargs;

static void ppmlr_w(win_t* W, const wargs* A) {
  const pn_params* P = A->P;
  const int nmax = A->nmax, axis = A->axis;
  const double dt = A->dt, hdt = 0.5 * A->dt;
  double *r = A->r, *p = A->p, *u = A->u, *v = A->v, *w = A->w, *e = A->e;
  double *xa0 = A->xa0, *dx0 = A->dx0;
  double gam = P->gam, gamm = P->gam - 1.0;
  int n;

  for (n = 0; n < nmax; n++) {
    W->q[n] = 0.0;
    W->dx[n] = 0.0;
    W->umid[n] = 0.0;
    W->pmid[n] = 0.0;
    W->tmp1[n] = 0.0;
    W->flat[n] = 0.0;
  }
  if (P->flat_flag) flatten(W, nmax, p, u, W->tmp1);

  /* ---- states (states.cu:3-73); al=q, da=dx, a6=tmp1, cs=umid, csfac=pmid */
  double *al = W->q, *da = W->dx, *a6 = W->tmp1, *cs = W->umid, *csfac = W->pmid;
  paraset(W, 0, nmax, dx0, al, da, a6);
  for (n = 2; n < nmax - 1; n++) {
    double tmp;
    if (axis == 0)
      get_fx(P, 1, n, hdt, xa0[n], A->azi[n] + hdt * A->v_rot[n], A->pol[n], &tmp, &W->ftot_s[n], A->pl, v, w);
    else if (axis == 1)
      get_fy(P, hdt, A->rad[n], xa0[n] + hdt * A->v_rot[n], A->pol[n], &tmp, &W->ftot_s[n], A->pl);
    else
      get_fz(P, 1, n, hdt, A->rad[n], A->azi[n] + hdt * A->v_rot[n], xa0[n], &tmp, &W->ftot_s[n], A->pl, w);
  }
  for (n = 0; n < nmax; n++) {
    cs[n] = hdt * sqrt(gam * p[n] / r[n]) / dx0[n];
    if (axis == 1) cs[n] /= A->rad[n];
    if (axis == 2 && P->ngeomz == 5) cs[n] /= A->rad[n];
    csfac[n] = 1.0 - fourthd * cs[n];
  }
  parabola(W, 0, nmax, W->flat, p, al, da, a6);
  for (n = 3; n < nmax - 2; n++) {
    double tmp = al[n - 1] + da[n - 1] - cs[n - 1] * (da[n - 1] - csfac[n - 1] * a6[n - 1]);
    W->Spl[n] = fmax(smallp, tmp);
    tmp = al[n] + cs[n] * (da[n] + csfac[n] * a6[n]);
    W->Spr[n] = fmax(smallp, tmp);
  }
  parabola(W, 0, nmax, W->flat, r, al, da, a6);
  for (n = 3; n < nmax - 2; n++) {
    double tmp = al[n - 1] + da[n - 1] - cs[n - 1] * (da[n - 1] - csfac[n - 1] * a6[n - 1]);
    W->Srl[n] = fmax(smallr, tmp);
    tmp = al[n] + cs[n] * (da[n] + csfac[n] * a6[n]);
    W->Srr[n] = fmax(smallr, tmp);
  }
  parabola(W, 0, nmax, W->flat, u, al, da, a6);
  for (n = 3; n < nmax - 2; n++) {
    double tmp = al[n - 1] + da[n - 1] - cs[n - 1] * (da[n - 1] - csfac[n - 1] * a6[n - 1]);
    W->Sul[n] = tmp + hdt * W->ftot_s[n];
    tmp = al[n] + cs[n] * (da[n] + csfac[n] * a6[n]);
    W->Sur[n] = tmp + hdt * W->ftot_s[n];
  }

  /* ---- riemann (riemann.cu:3-115) */
  for (n = 3; n < nmax - 2; n++)
    orc_riemann(P->eos, gam, W->Spl[n], W->Srl[n], W->Sul[n], W->Spr[n], W->Srr[n], W->Sur[n], &W->umid[n],
                &W->pmid[n]);

  /* ---- evolve (evolve.cu:3-166); vfac = tmp1 */
  double* vfac = W->tmp1;
  double *umid = W->umid, *pmid = W->pmid, *dx = W->dx, *q = W->q;
  for (n = 0; n < nmax; n++) {
    W->xa[n] = xa0[n];
    W->dvol[n] = A->dvol0[n];
    vfac[n] = xa0[n] + 0.5 * dx0[n];
  }
  for (n = 3; n < nmax - 2; n++) {
    double delx;
    if (axis == 0) {
      delx = dt * umid[n];
    } else if (axis == 1) {
      if (P->ngeomy > 2) delx = dt * (umid[n] / A->rad[n] - A->v_rot[n] - A->FrRot);
      else delx = dt * umid[n];
    } else {
      if (P->ngeomz == 5) delx = dt * umid[n] / A->rad[n];
      else delx = dt * umid[n];
    }
    W->dtdm[n] = dt / (r[n] * A->dvol0[n]);
    W->delx[n] = delx;
  }
  for (n = 3; n < nmax - 2; n++) vfac[n] = xa0[n] + W->delx[n];
  for (n = 3; n < nmax - 3; n++) {
    dx[n] = vfac[n + 1] - vfac[n];
    W->xa[n] = vfac[n];
  }
  for (n = 3; n < nmax - 3; n++) {
    double xa = W->xa[n];
    double xcl = xa + 0.5 * dx[n];
    W->xcl[n] = xcl;
    if (axis == 0) {
      get_fx(P, 0, n, 0.0, A->rad[n], A->azi[n], A->pol[n], &W->gra1[n], &W->fto1[n], A->pl, v, w);
      get_fx(P, 0, n, dt, xcl, A->azi[n] + dt * A->v_rot[n], A->pol[n], &W->gra2[n], &W->fto2[n], A->pl, v, w);
    } else if (axis == 1) {
      get_fy(P, 0.0, A->rad[n], A->azi[n], A->pol[n], &W->gra1[n], &W->fto1[n], A->pl);
      get_fy(P, dt, A->rad[n], xcl + dt * A->v_rot[n], A->pol[n], &W->gra2[n], &W->fto2[n], A->pl);
    } else {
      get_fz(P, 0, n, 0.0, A->rad[n], A->azi[n], A->pol[n], &W->gra1[n], &W->fto1[n], A->pl, w);
      get_fz(P, 0, n, dt, A->rad[n], A->azi[n] + dt * A->v_rot[n], xcl, &W->gra2[n], &W->fto2[n], A->pl, w);
    }
  }
  for (n = 3; n < nmax - 2; n++) {
    double xa = W->xa[n];
    umid[n] *= pmid[n];
    if (axis == 0) {
      if (P->ngeomx == 0) {
        W->dvol[n] = dx[n];
        vfac[n] = 1.0;
      } else if (P->ngeomx == 1) {
        W->dvol[n] = dx[n] * (xa + 0.5 * dx[n]);
        vfac[n] = 0.5 * (xa + xa0[n]);
      } else if (P->ngeomx == 2) {
        W->dvol[n] = dx[n] * (xa * (xa + dx[n]) + dx[n] * dx[n] * third);
        vfac[n] = (xa - xa0[n]) * (third * (xa - xa0[n]) + xa0[n]) + (xa0[n] * xa0[n]);
      }
    } else if (axis == 1) {
      if (P->ngeomy == 0) {
        W->dvol[n] = dx[n];
        vfac[n] = 1.0;
      } else {
        W->dvol[n] = dx[n] * A->rad[n];
        vfac[n] = 1.0;
      }
    } else {
      if (P->ngeomz == 0) {
        W->dvol[n] = dx[n];
        vfac[n] = 1.0;
      } else if (P->ngeomz == 5) {
        double delx = xa - xa0[n];
        double dum = cos(xa);
        W->dvol[n] = (dum - cos(xa + dx[n])) * A->rad[n];
        if (delx == 0.0) vfac[n] = sin(xa);
        else vfac[n] = (cos(xa0[n]) - dum) / delx;
      }
    }
  }
  for (n = 3; n < nmax - 3; n++) {
    if (P->eos == 0) q[n] = p[n] / r[n];
    else if (P->eos == 1) q[n] = p[n] / pow(r[n], gam);
    r[n] = fmax(r[n] * (A->dvol0[n] / W->dvol[n]), smallr);
    double dum = u[n];
    double delx = 0.5 * dt * (W->fto1[n] + W->fto2[n]);
    delx += -W->dtdm[n] * (pmid[n + 1] - pmid[n]) * 0.5 * (vfac[n + 1] + vfac[n]);
    u[n] = dum + delx;
    if (P->eos == 2) {
      double xcl = W->xcl[n];
      double ef = e[n] - W->dtdm[n] * (vfac[n + 1] * umid[n + 1] - vfac[n] * umid[n]);
      if (axis == 0)
        q[n] = ef - 0.5 * ((u[n] * u[n]) + (v[n] * v[n] / xcl / xcl)) - orc_potential(P, A->pl, xcl, A->azi[n], A->pol[n]);
      else if (axis == 1)
        q[n] = ef - 0.5 * (u[n] * u[n]) - orc_potential(P, A->pl, A->rad[n], xcl, A->pol[n]);
      else
        q[n] = ef - 0.5 * (u[n] * u[n]) - orc_potential(P, A->pl, A->rad[n], A->azi[n], xcl);
      q[n] = fmax(gamm * q[n], smallp);
    } else if (P->eos == 3) {
      /* e = p / (gamm rho) + u^2/2 with the pressure work and the time-centred work of the
         forces that moved u (the commented line of evolve.cu:146); p from what remains */
      double ef = e[n] - W->dtdm[n] * (vfac[n + 1] * umid[n + 1] - vfac[n] * umid[n]) +
                  0.5 * dt * (dum * W->fto1[n] + u[n] * W->fto2[n]);
      q[n] = fmax(gamm * (ef - 0.5 * (u[n] * u[n])), smallp);
    }
  }

  /* ---- remap (remap.cu:14-226); al=tmp1, da=umid, a6=pmid */
  al = W->tmp1;
  da = W->umid;
  a6 = W->pmid;
  paraset(W, 3, nmax - 3, dx, al, da, a6);
  for (n = 0; n < nmax; n++) {
    double xa = W->xa[n], x0 = xa0[n], lvol = 0.0;
    if (axis == 0) {
      if (P->ngeomx == 0) lvol = xa - x0;
      else if (P->ngeomx == 1) {
        lvol = xa - x0;
        lvol *= (x0 + 0.5 * lvol);
      } else if (P->ngeomx == 2) {
        lvol = xa - x0;
        lvol *= x0 * xa + third * (lvol * lvol);
      }
    } else if (axis == 1) {
      if (P->ngeomy == 0) lvol = xa - x0;
      else lvol = (xa - x0) * A->rad[n];
    } else {
      if (P->ngeomz == 0) lvol = xa - x0;
      else if (P->ngeomz == 5) lvol = (cos(x0) - cos(xa)) * A->rad[n];
    }
    W->lvol[n] = lvol;
    W->dxf[n] = xa - x0;
    W->delC[n] = (W->dxf[n] >= 0.0);
  }
  parabola(W, 3, nmax - 3, W->flat, r, al, da, a6);
  for (n = 6; n < nmax - 5; n++) {
    W->dxn[n] = dx[n];
    W->dxnm[n] = dx[n - 1];
  }
  double* flux = dx;
  for (n = 6; n < nmax - 5; n++) {
    if (W->delC[n]) {
      W->dxf[n] = 0.5 * W->dxf[n] / W->dxnm[n];
      W->dxf2[n] = 1.0 - fourthd * W->dxf[n];
      W->fr[n] = (al[n - 1] + da[n - 1] - W->dxf[n] * (da[n - 1] - W->dxf2[n] * a6[n - 1])) * W->lvol[n];
    } else {
      W->dxf[n] = 0.5 * W->dxf[n] / W->dxn[n];
      W->dxf2[n] = 1.0 + fourthd * W->dxf[n];
      W->fr[n] = (al[n] - W->dxf[n] * (da[n] + W->dxf2[n] * a6[n])) * W->lvol[n];
    }
  }
  for (n = 6; n < nmax - 5; n++) flux[n] = W->fr[n];
  for (n = 6; n < nmax - 6; n++) {
    W->m[n] = r[n] * W->dvol[n];
    double lv = (W->m[n] + W->fr[n] - flux[n + 1]) / A->dvol0[n];
    r[n] = fmax(smallr, lv);
    W->invm[n] = 1.0 / (r[n] * A->dvol0[n]);
  }
  /* u, v, w and q: parabola, face flux = face value * mass flux, update */
  double* fld[4] = {u, v, w, q};
  int nf = 4; /* q is remapped for every EOS (remap.cu:144-176) */
  for (int f = 0; f < nf; f++) {
    double* x = fld[f];
    parabola(W, 3, nmax - 3, W->flat, x, al, da, a6);
    double fl[ARR + 1];
    for (n = 6; n < nmax - 5; n++) {
      if (W->delC[n]) flux[n] = (al[n - 1] + da[n - 1] - W->dxf[n] * (da[n - 1] - W->dxf2[n] * a6[n - 1])) * W->fr[n];
      else flux[n] = (al[n] - W->dxf[n] * (da[n] + W->dxf2[n] * a6[n])) * W->fr[n];
    }
    for (n = 6; n < nmax - 6; n++) fl[n] = flux[n] - flux[n + 1];
    for (n = 6; n < nmax - 6; n++) x[n] = (x[n] * W->m[n] + fl[n]) * W->invm[n];
    if (f == 3) {
      for (n = 6; n < nmax - 6; n++) {
        if (P->eos == 1) p[n] = fmax(smallp, pow(r[n], gam) * q[n]);
        else p[n] = fmax(smallp, r[n] * q[n]);
      }
    }
  }
}

/* derivs, viscosity.cu:2-19 */
void orc_derivs(const double* xf, const double* u, int n, double* first, double* second) {
  double x2 = xf[n] - xf[n - 1];
  double x3 = xf[n + 1] - xf[n - 1];
  double h2 = (u[n] - u[n - 1]) / x2;
  double h3 = (u[n + 1] - u[n - 1]) / x3;
  double A = (h3 - h2) / (x3 - x2);
  double B = (x3 * h2 - x2 * h3) / (x3 - x2);
  *second = 2.0 * A;
  *first = *second * x2 + B;
}

/* device_viscosity_r, viscosity.cu:370-421; R per cell */
static void visc_r(win_t* W, const pn_params* P, int lim, double dt, const double* R, const double* x,
                   const double* cell_r, double* u, double* v, double* w, double* nn) {
  for (int i = 0; i < lim; i++) {
    W->rho[i] = cell_r[i];
    W->nr[i] = get_nu(P, R[i]) * W->rho[i];
    nn[i] = W->nr[i];
    v[i] /= R[i] * R[i];
  }
  for (int i = NPAD; i < lim - NPAD; i++) {
    double dn, d1, d2, nr = W->nr[i], rho = W->rho[i], Ri = R[i];
    orc_derivs(x, nn, i, &dn, &d2);
    orc_derivs(x, u, i, &d1, &d2);
    double ddu = 2.0 * nr * d2 + 2.0 * dn * d1 + 2.0 * nr * d1 / Ri - 2.0 * nr * u[i] / Ri / Ri;
    W->ddu[i] = u[i] + dt * ddu / rho;
    orc_derivs(x, v, i, &d1, &d2);
    double ddv = Ri * nr * d2 + Ri * dn * d1 + 3.0 * nr * d1;
    W->ddv[i] = v[i] * Ri + dt * ddv / rho;
    if (P->ndim == 3) {
      orc_derivs(x, w, i, &d1, &d2);
      double ddw = nr * d2 + dn * d1 + nr * d1 / Ri;
      W->ddw[i] = w[i] + dt * ddw / rho;
    }
  }
  for (int i = NPAD; i < lim - NPAD; i++) {
    u[i] = W->ddu[i];
    v[i] = W->ddv[i] * R[i];
    if (P->ndim == 3) w[i] = W->ddw[i];
  }
}

/* device_viscosity_p, viscosity.cu:423-472 */
static void visc_p(win_t* W, const pn_params* P, int lim, double dt, double R, const double* x,
                   const double* cell_r, double* u, double* v, double* w, double* nn) {
  for (int j = 0; j < lim; j++) {
    W->rho[j] = cell_r[j];
    W->nr[j] = get_nu(P, R) * W->rho[j];
    nn[j] = W->nr[j];
  }
  for (int j = NPAD; j < lim - NPAD; j++) {
    double dn, d1, d2, nr = W->nr[j], rho = W->rho[j];
    orc_derivs(x, nn, j, &dn, &d2);
    orc_derivs(x, v, j, &d1, &d2);
    double ddv = 2.0 * nr * d2 + 2.0 * dn * d1;
    double ddu = 0.0;
    W->ddv[j] = v[j] + dt * ddv / rho;
    if (P->ndim == 3) {
      orc_derivs(x, w, j, &d1, &d2);
      double ddw = nr * d2 + dn * d1;
      W->ddw[j] = w[j] + dt * ddw / rho;
    }
    orc_derivs(x, u, j, &d1, &d2);
    ddu += nr * d2 + dn * d1;
    W->ddu[j] = u[j] + dt * ddu / rho;
  }
  for (int j = NPAD; j < lim - NPAD; j++) {
    v[j] = W->ddv[j];
    if (P->ndim == 3) w[j] = W->ddw[j];
    u[j] = W->ddu[j];
  }
}

/* device_viscosity_z, viscosity.cu:474-518: nu = vis_nu (no radial law) times rho; the
   component along the pencil (w of the reference) takes 2 nu rho d2 + 2 dn d1, the two
   transverse ones nu rho d2 + dn d1; no curvature terms.  The reference has no caller; the
   theta sweep calls it with (u, v, w) = (v_r, v_phi, v_theta) and x = r times the theta left
   edges (arc length, as the phi sweep passes r_cyl times its edges, ppm.cu:277-278).  The
   reference updates 0 < k < arrsize - 1; only the stored cells [NPAD, lim - NPAD) matter. */
static void visc_z(win_t* W, const pn_params* P, int lim, double dt, const double* x, const double* cell_r,
                   double* u, double* v, double* w, double* nn) {
  for (int k = 0; k < lim; k++) {
    W->rho[k] = cell_r[k];
    W->nr[k] = P->vis_nu * W->rho[k];
    nn[k] = W->nr[k];
  }
  for (int k = NPAD; k < lim - NPAD; k++) {
    double dn, d1, d2, nr = W->nr[k], rho = W->rho[k];
    orc_derivs(x, nn, k, &dn, &d2);
    orc_derivs(x, w, k, &d1, &d2);
    double ddw = 2.0 * nr * d2 + 2.0 * dn * d1;
    orc_derivs(x, u, k, &d1, &d2);
    double ddu = nr * d2 + dn * d1;
    orc_derivs(x, v, k, &d1, &d2);
    double ddv = nr * d2 + dn * d1;
    W->ddw[k] = w[k] + dt * ddw / rho;
    W->ddu[k] = u[k] + dt * ddu / rho;
    W->ddv[k] = v[k] + dt * ddv / rho;
  }
  for (int k = NPAD; k < lim - NPAD; k++) {
    w[k] = W->ddw[k];
    u[k] = W->ddu[k];
    v[k] = W->ddv[k];
  }
}

/* ------------------------------------------------------------------------ */
/* exposed building blocks                                                   */

void orc_paraset_coeffs(int nmax, int nmin, int nmax_p, const double* dx, double* c) {
  win_t* W = calloc(1, sizeof(win_t));
  double a[ARR], ai[ARR], b[ARR];
  (void)nmax;
  paraset(W, nmin, nmax_p, dx, a, ai, b);
  for (int n = 0; n < nmax && n < ARR; n++) {
    c[5 * n + 0] = W->c0[n];
    c[5 * n + 1] = W->c1[n];
    c[5 * n + 2] = W->c2[n];
    c[5 * n + 3] = W->c3[n];
    c[5 * n + 4] = W->c4[n];
  }
  free(W);
}

void orc_parabola_window(int nmax, int nmin, int nmax_p, const double* dx, const double* val,
                         const double* flat, double* al, double* da, double* a6) {
  win_t* W = calloc(1, sizeof(win_t));
  double a[ARR], ai[ARR], b[ARR];
  (void)nmax;
  paraset(W, nmin, nmax_p, dx, a, ai, b);
  parabola(W, nmin, nmax_p, flat, val, al, da, a6);
  free(W);
}

void orc_flatten_window(int nmax, const double* p, const double* u, double* flat) {
  win_t* W = calloc(1, sizeof(win_t));
  double steep[ARR];
  flatten(W, nmax, p, u, steep);
  for (int n = 0; n < nmax; n++) flat[n] = W->flat[n];
  free(W);
}

static void window_update(win_t* W, const pn_params* P, int axis, int nmax, double dt, double FrRot,
                          const pn_body* b, const double* rad, const double* azi, const double* pol,
                          const double* v_rot, const double* dvol0, double* xa0, double* dx0, double* r,
                          double* p, double* u, double* v, double* w, double* e) {
  wargs A = {P, axis, nmax, dt, FrRot, b, rad, azi, pol, v_rot, dvol0, xa0, dx0, r, p, u, v, w, e};
  ppmlr_w(W, &A);
}

void orc_ppmlr_window(const pn_params* P, int axis, int nmax, double dt, double FrRot, const pn_body* b,
                      const double* rad, const double* azi, const double* pol, const double* v_rot,
                      const double* dvol0, double* xa0, double* dx0, double* r, double* p, double* u,
                      double* v, double* w) {
  win_t* W = calloc(1, sizeof(win_t));
  double e[ARR];
  memset(e, 0, sizeof e);
  window_update(W, P, axis, nmax, dt, FrRot, b, rad, azi, pol, v_rot, dvol0, xa0, dx0, r, p, u, v, w, e);
  free(W);
}

/* ------------------------------------------------------------------------ */
/* sweeps                                                                    */

/* number of windows and the length of the last one for a pencil of len
   cells; a length that is a multiple of 116 gets a full last window */
static int nwin(int len) { return (len + REAL - 1) / REAL; }
static int win_lim(int len, int l) {
  int L = nwin(len);
  if (l < L - 1) return ARR;
  return len - REAL * (L - 1) + 2 * NPAD;
}

static double eos_store_p(const pn_params* P, double r, double p, double rad_cyl) {
  if (P->eos == 0) return r * get_cs2(P, rad_cyl);
  if (P->eos == 1) return get_cs2(P, rad_cyl) * pow(r, P->gam) / P->gam;
  return p;
}

/* sweepx, ppm.cu:26-160: one radial pencil per lab column j0 and theta ring k */
/* set_boundx (boundary.cu:86-206): the values a radial ghost cell of theta row k at lab
   column j0 takes under the boundary codes; v[0..4] = r, p, u, v, w arrive holding the fixed
   ghost ring's values (code 3).  Ghost ring i_ring < 0 lies t = -i_ring cells outside the
   left edge cell, i_ring >= imax lies t = i_ring - imax + 1 cells outside the right one. */
static void ghost_values(const pn_grid* g, const pn_params* P, const void* rings, int k, int j0, int i_ring,
                         double* v) {
  const int imax = g->imax, jmax = g->jmax;
  const int left = i_ring < 0;
  const int code = left ? P->nlft : P->nrgh;
  const int t = left ? -i_ring : i_ring - imax + 1;
  int src, first = 0, sign_u = 1;
  switch (code) {
    case 0: src = left ? t : imax - 1 - t; sign_u = -1; break;  /* reflect about the edge cell */
    case 1: src = left ? 0 : imax - 1; break;                    /* outflow */
    case 2: src = left ? imax - t : t - 1; break;                /* periodic */
    case 4: src = left ? 0 : imax - 1; break;                    /* outflow of r, p (, u) */
    default: return;                                             /* 3: fixed ghost rings */
  }
  const int last = (code == 4) ? (left ? 3 : 2) : 5;   /* fields [first, last) come from src */
  ring_t S = ring_at(rings, jmax, src + (long)imax * k);
  int j = j0 - *S.rot_j;
  if (j < 0) j += jmax;
  if (j >= jmax) j -= jmax;
  const double sv[5] = {S.r[j], S.p[j], S.u[j], S.v[j], S.w[j]};
  for (int f = first; f < last; f++) v[f] = sv[f];
  if (code == 0) v[2] = sign_u * sv[2];
}

void orc_sweepx(const pn_grid* g, const pn_params* P, void* rings, const void* lft, const void* rgh,
                double dt, const pn_body* planet, double FrRot) {
  const int imax = g->imax, jmax = g->jmax, kmax = g->kmax;
  const int L = nwin(imax);
#pragma omp parallel
  {
    win_t* W = calloc(1, sizeof(win_t));
    int plen = imax + 2 * NPAD;
    double* pr = malloc(sizeof(double) * 5 * plen);
    ring_t* cur = malloc(sizeof(ring_t) * plen);
    int* jd = malloc(sizeof(int) * plen);
#pragma omp for collapse(2) schedule(static)
    for (int k = 0; k < kmax; k++)
      for (int j0 = 0; j0 < jmax; j0++) {
        /* gather the pre-update pencil (the tile carry of ppm.cu:88-118
           hands each window pre-update halo values) */
        for (int c = 0; c < plen; c++) {
          int i_ring = c - NPAD;
          ring_t R;
          if (i_ring < 0) R = ring_at(lft, jmax, i_ring + NPAD + NPAD * k);
          else if (i_ring >= imax) R = ring_at(rgh, jmax, i_ring - imax + NPAD * k);
          else R = ring_at(rings, jmax, i_ring + (long)imax * k);
          cur[c] = R;
          int j = j0 - *R.rot_j;
          if (j < 0) j += jmax;
          if (j >= jmax) j -= jmax;
          jd[c] = j;
          pr[5 * c + 0] = R.r[j];
          pr[5 * c + 1] = R.p[j];
          pr[5 * c + 2] = R.u[j];
          pr[5 * c + 3] = R.v[j];
          pr[5 * c + 4] = R.w[j];
          if (i_ring < 0 || i_ring >= imax) ghost_values(g, P, rings, k, j0, i_ring, pr + 5 * c);
        }
        for (int l = 0; l < L; l++) {
          int lim = win_lim(imax, l);
          double r[ARR], p[ARR], u[ARR], v[ARR], w[ARR], e[ARR], xa0[ARR], dx0[ARR];
          double rad[ARR], azi[ARR], pol[ARR], vrot[ARR], dvol0[ARR], radc[ARR];
          for (int i = 0; i < lim; i++) {
            int c = REAL * l + i;
            ring_t R = cur[c];
            rad[i] = (P->ngeomx > 0) ? *R.xc : 1.0;
            azi[i] = R.yc[j0];
            pol[i] = (P->ndim == 3) ? *R.zc : 0.0;
            radc[i] = rad[i];
            if (P->ngeomz == 5) radc[i] *= sin(pol[i]);
            xa0[i] = *R.x;
            dx0[i] = *R.dx;
            dvol0[i] = *R.xvol;
            r[i] = pr[5 * c + 0];
            p[i] = pr[5 * c + 1];
            u[i] = pr[5 * c + 2];
            v[i] = pr[5 * c + 3];
            w[i] = pr[5 * c + 4];
            if (P->ngeomy > 2) v[i] *= radc[i];
            if (P->ngeomz == 5) w[i] *= rad[i];
            e[i] = orc_energy(P, p[i], r[i], u[i], v[i]);
            vrot[i] = *R.res_v / radc[i];
          }
          window_update(W, P, 0, lim, dt, FrRot, planet, rad, azi, pol, vrot, dvol0, xa0, dx0, r, p, u, v, w, e);
          if (P->visc_flag == 1) visc_r(W, P, lim, dt, radc, xa0, r, u, v, w, dx0);
          for (int i = NPAD; i < lim - NPAD; i++) {
            int c = REAL * l + i;
            ring_t R = cur[c];
            int j = jd[c];
            R.r[j] = r[i];
            R.p[j] = eos_store_p(P, r[i], p[i], radc[i]);
            R.u[j] = u[i];
            R.v[j] = (P->ngeomy > 2) ? v[i] / radc[i] : v[i];
            R.w[j] = (P->ngeomz == 5) ? w[i] / rad[i] : w[i];
          }
        }
      }
    free(W);
    free(pr);
    free(cur);
    free(jd);
  }
}

/* sweepy, ppm.cu:162-307: one periodic azimuthal pencil per ring */
void orc_sweepy(const pn_grid* g, const pn_params* P, void* rings, double dt, const pn_body* planet,
                double FrRot) {
  const int imax = g->imax, jmax = g->jmax, kmax = g->kmax;
  const int L = nwin(jmax);
  const long nring = (long)imax * kmax;
#pragma omp parallel
  {
    win_t* W = calloc(1, sizeof(win_t));
    double* pr = malloc(sizeof(double) * 5 * jmax);
#pragma omp for schedule(static)
    for (long nrg = 0; nrg < nring; nrg++) {
      ring_t R = ring_at(rings, jmax, nrg);
      double rad = (P->ngeomx > 0) ? *R.xc : 1.0;
      double pol = (P->ndim == 3) ? *R.zc : 0.0;
      double rad_cyl = rad;
      if (P->ngeomz == 5) rad_cyl *= sin(pol);
      int rot = *R.rot_j;
      /* pre-update pencil in lab order, with the component swap u<-v, v<-w, w<-u */
      for (int j0 = 0; j0 < jmax; j0++) {
        int j = j0 - rot;
        if (j < 0) j += jmax;
        if (j >= jmax) j -= jmax;
        pr[5 * j0 + 0] = R.r[j];
        pr[5 * j0 + 1] = R.p[j];
        pr[5 * j0 + 2] = R.v[j];
        pr[5 * j0 + 3] = R.w[j];
        pr[5 * j0 + 4] = R.u[j];
      }
      for (int l = 0; l < L; l++) {
        int lim = win_lim(jmax, l);
        double r[ARR], p[ARR], u[ARR], v[ARR], w[ARR], e[ARR], xa0[ARR], dx0[ARR];
        double radv[ARR], azi[ARR], polv[ARR], vrot[ARR], dvol0[ARR];
        for (int i = 0; i < lim; i++) {
          int jref = REAL * l + i - NPAD;
          int j0 = jref;
          if (j0 < 0) j0 += jmax;
          if (j0 >= jmax) j0 -= jmax;
          azi[i] = R.yc[j0];
          xa0[i] = R.y[j0];
          dx0[i] = R.dy[j0];
          dvol0[i] = R.yvol[j0];
          if (jref < 0) {
            xa0[i] -= twopi;
            azi[i] -= twopi;
          }
          if (jref >= jmax) {
            xa0[i] += twopi;
            azi[i] += twopi;
          }
          if (P->ngeomy > 2) dvol0[i] *= rad_cyl;
          r[i] = pr[5 * j0 + 0];
          p[i] = pr[5 * j0 + 1];
          u[i] = pr[5 * j0 + 2];
          v[i] = pr[5 * j0 + 3];
          w[i] = pr[5 * j0 + 4];
          e[i] = orc_energy(P, p[i], r[i], u[i], v[i]);
          radv[i] = rad_cyl;
          polv[i] = pol;
          vrot[i] = *R.res_v / rad_cyl;
        }
        window_update(W, P, 1, lim, dt, FrRot, planet, radv, azi, polv, vrot, dvol0, xa0, dx0, r, p, u, v, w, e);
        if (P->visc_flag == 1) {
          for (int i = 0; i < lim; i++) xa0[i] *= rad_cyl;
          visc_p(W, P, lim, dt, rad_cyl, xa0, r, w, u, v, dx0);
        }
        for (int i = NPAD; i < lim - NPAD; i++) {
          int j0 = REAL * l + i - NPAD;
          int j = j0 - rot;
          if (j < 0) j += jmax;
          if (j >= jmax) j -= jmax;
          R.r[j] = r[i];
          R.p[j] = eos_store_p(P, r[i], p[i], rad_cyl);
          R.u[j] = w[i];
          R.v[j] = u[i];
          R.w[j] = v[i];
        }
      }
    }
    free(W);
    free(pr);
  }
}

/* theta sweep (3-D).  The reference has no caller for it (main.cu:376 calls
   an undefined bundle_sweep3D); it is assembled from the theta pieces the
   reference does hold: get_fz_bound / get_fz (forces.cu:298-372), the theta
   branches of states / evolve / remap (states.cu:29-39, evolve.cu:26-29,
   63-67, 110-124, remap.cu:50-54) and the udr / top ghost rings
   (init.cpp:415-541).  Pencil: one per (i, lab column j0) over theta rings
   k, ghost rings below / above, component swap u<-w (v_theta),
   v<-u (v_r), w<-v*rad_cyl (angular momentum about the axis, the
   form get_fz's centrifugal term expects).  With viscosity, the window then
   takes device_viscosity_z (viscosity.cu:474-518) on (v_r, v_phi, v_theta). */
void orc_sweepz(const pn_grid* g, const pn_params* P, void* rings, const void* udr, const void* top,
                double dt, const pn_body* planet, double FrRot) {
  const int imax = g->imax, jmax = g->jmax, kmax = g->kmax;
  const int L = nwin(kmax);
  (void)FrRot;
#pragma omp parallel
  {
    win_t* W = calloc(1, sizeof(win_t));
    int plen = kmax + 2 * NPAD;
    double* pr = malloc(sizeof(double) * 5 * plen);
    ring_t* cur = malloc(sizeof(ring_t) * plen);
    int* jd = malloc(sizeof(int) * plen);
#pragma omp for collapse(2) schedule(static)
    for (int i = 0; i < imax; i++)
      for (int j0 = 0; j0 < jmax; j0++) {
        for (int c = 0; c < plen; c++) {
          int kk = c - NPAD;
          ring_t R;
          if (kk < 0) R = ring_at(udr, jmax, kk + NPAD + NPAD * i);
          else if (kk >= kmax) R = ring_at(top, jmax, kk - kmax + NPAD * i);
          else R = ring_at(rings, jmax, i + (long)imax * kk);
          cur[c] = R;
          int j = j0 - *R.rot_j;
          if (j < 0) j += jmax;
          if (j >= jmax) j -= jmax;
          jd[c] = j;
          pr[5 * c + 0] = R.r[j];
          pr[5 * c + 1] = R.p[j];
          pr[5 * c + 2] = R.u[j];
          pr[5 * c + 3] = R.v[j];
          pr[5 * c + 4] = R.w[j];
        }
        for (int l = 0; l < L; l++) {
          int lim = win_lim(kmax, l);
          double r[ARR], p[ARR], u[ARR], v[ARR], w[ARR], e[ARR], xa0[ARR], dx0[ARR];
          double rad[ARR], azi[ARR], pol[ARR], vrot[ARR], dvol0[ARR], radc[ARR];
          for (int kk = 0; kk < lim; kk++) {
            int c = REAL * l + kk;
            ring_t R = cur[c];
            rad[kk] = (P->ngeomx > 0) ? *R.xc : 1.0;
            azi[kk] = R.yc[j0];
            pol[kk] = *R.zc;
            radc[kk] = rad[kk];
            if (P->ngeomz == 5) radc[kk] *= sin(pol[kk]);
            xa0[kk] = *R.z;
            dx0[kk] = *R.dz;
            dvol0[kk] = *R.zvol;
            if (P->ngeomz == 5) dvol0[kk] *= rad[kk];
            r[kk] = pr[5 * c + 0];
            p[kk] = pr[5 * c + 1];
            u[kk] = pr[5 * c + 4];
            v[kk] = pr[5 * c + 2];
            w[kk] = pr[5 * c + 3] * radc[kk];
            e[kk] = orc_energy(P, p[kk], r[kk], u[kk], v[kk]);
            vrot[kk] = *R.res_v / radc[kk];
          }
          window_update(W, P, 2, lim, dt, FrRot, planet, rad, azi, pol, vrot, dvol0, xa0, dx0, r, p, u, v, w, e);
          /* v_phi from the angular momentum the sweep carried */
          double vphi[ARR];
          for (int kk = 0; kk < lim; kk++) vphi[kk] = w[kk] / radc[kk];
          if (P->visc_flag == 1) {
            for (int kk = 0; kk < lim; kk++) xa0[kk] *= rad[kk];
            visc_z(W, P, lim, dt, xa0, r, v, vphi, u, dx0);
          }
          for (int kk = NPAD; kk < lim - NPAD; kk++) {
            int c = REAL * l + kk;
            ring_t R = cur[c];
            int j = jd[c];
            R.r[j] = r[kk];
            R.p[j] = eos_store_p(P, r[kk], p[kk], radc[kk]);
            R.w[j] = u[kk];
            R.u[j] = v[kk];
            R.v[j] = vphi[kk];
          }
        }
      }
    free(W);
    free(pr);
    free(cur);
    free(jd);
  }
}

/* rotate_rings, ppm.cu:309-322 */
void orc_rotate(const pn_grid* g, void* rings) {
  long nring = (long)g->imax * g->kmax;
  for (long n = 0; n < nring; n++) {
    ring_t R = ring_at(rings, g->jmax, n);
    int j = *R.rot_j + *R.inc_j;
    if (j < 0) j += g->jmax;
    if (j >= g->jmax) j -= g->jmax;
    *R.rot_j = j;
  }
}

/* Index_Shift, orbital_advection.cu:12-43 */
void orc_index_shift(const pn_grid* g, const pn_params* P, void* rings, const pn_symdisk* val, double dt,
                     double FrRot) {
  long nring = (long)g->imax * g->kmax;
  for (long n = 0; n < nring; n++) {
    ring_t R = ring_at(rings, g->jmax, n);
    double widthy = *R.xc * R.dy[0] / dt;
    double rot_v;
    if (P->fargo_flag == 1) {
      int idx = *R.i + g->imax * *R.k;
      rot_v = val[idx].v - FrRot * *R.xc;
    } else if (P->fargo_flag == 2) {
      rot_v = (1.0 - FrRot) * *R.xc;
    } else {
      rot_v = 0.0;
    }
    double jp = (*R.rot_v / widthy);
    if (jp >= 0.0) {
      *R.inc_j = (int)(jp + 0.5);
      *R.res_v = floor(jp + 0.5) * widthy;
    } else {
      *R.inc_j = (int)(jp - 0.5);
      *R.res_v = floor(jp + 0.5) * widthy;
    }
    *R.rot_v = rot_v;
  }
}

/* get_dt, timestep.cu:12-149 (every ring reduced) */
double orc_get_dt(const pn_grid* g, const pn_params* P, const void* rings, double dt_prev, double FrRot) {
  long nring = (long)g->imax * g->kmax;
  double mx = 0.0;
  for (long n = 0; n < nring; n++) {
    ring_t R = ring_at(rings, g->jmax, n);
    double rad = *R.xc, rad_cyl = rad;
    if (P->ngeomz == 5) rad_cyl *= sin(*R.zc);
    for (int j = 0; j < g->jmax; j++) {
      double widthz = *R.dz, widthy = R.dy[j], widthx = *R.dx;
      if (P->ngeomy > 2) widthy *= rad_cyl;
      if (P->ngeomz > 2) widthz *= rad;
      double cs = sqrt(P->gam * R.p[j] / R.r[j]);
      double t0 = (fabs(R.u[j]) + cs) / widthx;
      double t1 = (fabs(R.v[j] - *R.rot_v - FrRot * rad_cyl) + cs) / widthy;
      double val = fmax(t0, t1);
      if (P->ndim == 3) {
        double t2 = (fabs(R.w[j]) + cs) / widthz;
        val = max3(t0, t1, t2);
      }
      if (val > mx) mx = val;
    }
  }
  double dt = 1.1 * dt_prev;
  double dtc = P->courant / mx;
  return fmin(dt, dtc);
}

/* output_lv1 (device_func.cu:49-198), summed with Neumaier compensation */
double orc_reduce(const pn_grid* g, const pn_params* P, const void* rings, const pn_body* planet,
                  double exclude, int mode) {
  long nring = (long)g->imax * g->kmax;
  double s = 0.0, c = 0.0;
  exclude *= P->sc_h;
  for (long n = 0; n < nring; n++) {
    ring_t R = ring_at(rings, g->jmax, n);
    for (int j = 0; j < g->jmax; j++) {
      int j0 = j + *R.rot_j;
      if (j0 < 0) j0 += g->jmax;
      if (j0 >= g->jmax) j0 -= g->jmax;
      double rad = *R.xc, phi = R.yc[j0], tmp = 0.0;
      if (mode == 0 || P->plnt_flag == 0) {
        tmp = R.r[j] * *R.xvol * R.yvol[j0] * *R.zvol;
      } else if (P->ndim == 2 && (mode == 8 || mode == 9)) {
        double dd = (rad - planet->x) * (rad - planet->x) + rad * rad * (phi - planet->y) * (phi - planet->y);
        if (dd > exclude * exclude) {
          if (mode == 8) {
            tmp = -grav_rad_cyl(planet, P->bary_flag, rad, phi, 0.0, 0.0);
            tmp *= (R.r[j] - pow(*R.xc, -P->p_alpha)) * *R.xvol * R.yvol[j0];
          } else {
            tmp = -grav_azi_cyl(planet, P->bary_flag, rad, phi, 0.0, 0.0);
            tmp *= rad * (R.r[j] - pow(*R.xc, -P->p_alpha)) * *R.xvol * R.yvol[j0];
          }
        }
      }
      double t = s + tmp;
      if (fabs(s) >= fabs(tmp)) c += (s - t) + tmp;
      else c += (tmp - t) + s;
      s = t;
    }
  }
  return s + c;
}

/* total mass and angular momentum (the latter is not in the reference) */
void orc_totals(const pn_grid* g, const pn_params* P, const void* rings, double* mass, double* angmom) {
  long nring = (long)g->imax * g->kmax;
  double sm = 0.0, cm = 0.0, sl = 0.0, cl = 0.0;
  for (long n = 0; n < nring; n++) {
    ring_t R = ring_at(rings, g->jmax, n);
    double rad_cyl = *R.xc;
    if (P->ngeomz == 5) rad_cyl *= sin(*R.zc);
    for (int j = 0; j < g->jmax; j++) {
      int j0 = j + *R.rot_j;
      if (j0 >= g->jmax) j0 -= g->jmax;
      double dm = R.r[j] * *R.xvol * R.yvol[j0] * *R.zvol;
      double dl = dm * R.v[j] * rad_cyl;
      double t = sm + dm;
      if (fabs(sm) >= fabs(dm)) cm += (sm - t) + dm;
      else cm += (dm - t) + sm;
      sm = t;
      t = sl + dl;
      if (fabs(sl) >= fabs(dl)) cl += (sl - t) + dl;
      else cl += (dl - t) + sl;
      sl = t;
    }
  }
  *mass = sm + cm;
  *angmom = sl + cl;
}

/* The reader's ring average (write_output_file, read/p_output.cpp:112-127): over the lab
 * columns j of a ring, in order, sum f(lab j) dy[j] with f(lab j) at data index
 * (j - rot_j) mod jmax (update_grid, read/p_init.cpp:108-132), divided by the sum of dy[j]
 * (the reader divides by 2 pi, which that sum is on a full ring).  field 0..4 = r, p, u, v, w.
 * out[n] for ring n = i + imax k. */
void orc_ring_means(const pn_grid* g, const void* rings, int field, double* out) {
  long nring = (long)g->imax * g->kmax;
  for (long n = 0; n < nring; n++) {
    ring_t R = ring_at(rings, g->jmax, n);
    const double* F = field == 0 ? R.r : field == 1 ? R.p : field == 2 ? R.u : field == 3 ? R.v : R.w;
    double a = 0.0, w = 0.0;
    for (int j = 0; j < g->jmax; j++) {
      int jd = j - *R.rot_j;
      if (jd < 0) jd += g->jmax;
      a += F[jd] * R.dy[j];
      w += R.dy[j];
    }
    out[n] = a / w;
  }
}

/* OpenMP threads of the sweeps (the CPU baseline is timed on all cores and on one) */
void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* planet.cu:11-93                                                           */

static double set_M_p(const pn_params* P, double t) {
  if (t < P->t_growth) {
    double m = P->M_p * sin(pi_ * t / (2.0 * P->t_growth));
    return fmax(m, 0.01 * EarthMass);
  }
  return P->M_p;
}

void orc_init_planet(const pn_params* P, pn_body* b, double* FrRot) {
  memset(b, 0, sizeof *b);
  b->m = set_M_p(P, 0.0);
  b->x = P->R_p;
  b->y = pi_;
  b->vx = 0.0;
  b->vy = pow(b->x, -1.5);
  *FrRot = (P->frrot_flag == 1) ? b->vy : 0.0;
  b->vy -= *FrRot;
  b->rs = P->rs_fac * get_h(P, b->x);
}

void orc_planet_forces(const pn_grid* g, const pn_params* P, const void* rings, pn_body* b, double FrRot,
                       double simtime) {
  b->m = set_M_p(P, simtime);
  if (P->plnt_flag == 2) {
    double Sigma_0 = 1.0 * 0.00019126835;
    b->fx = Sigma_0 * orc_reduce(g, P, rings, b, 0.0, 8) / b->m - 1.0 / b->x / b->x + b->x * pow(b->vy + FrRot, 2);
    b->fy = Sigma_0 * orc_reduce(g, P, rings, b, 0.0, 9) / b->m;
  } else {
    b->fx = -1.0 / b->x / b->x + b->x * pow(b->vy + FrRot, 2);
    b->fy = 0.0;
  }
}

void orc_kick_drift_kick(const pn_params* P, pn_body* b, double* FrRot, double dt, int kick) {
  if (!kick) {
    double old_rad = b->x;
    b->x += dt * b->vx;
    b->y += dt * b->vy;
    b->vy = ((b->vy + *FrRot) * old_rad * old_rad / b->x / b->x) - *FrRot;
    if (b->y < 0.0) b->y += twopi;
    if (b->y >= twopi) b->y -= twopi;
    b->rs = P->rs_fac * get_h(P, b->x);
  } else {
    b->vx += dt * b->fx;
    b->vy += *FrRot + dt * b->fy / b->x / b->x;
    *FrRot = (P->frrot_flag == 1) ? b->vy : 0.0;
    b->vy -= *FrRot;
  }
}

/* ------------------------------------------------------------------------ */
/* step loop, main.cu:270-387                                                */

void orc_run_begin(const pn_grid* g, const pn_params* P, void* rings, const pn_symdisk* val, pn_run_state* st) {
  memset(st, 0, sizeof *st);
  st->dt = P->endtime;
  st->simtime = 0.0;
  orc_init_planet(P, &st->planet, &st->FrRot);
  orc_planet_forces(g, P, rings, &st->planet, st->FrRot, st->simtime);
  orc_index_shift(g, P, rings, val, st->dt, st->FrRot);
}

/* kill_wave (kill_wave.cu:3-43): cells within sc_h/2 of xmin or xmax relax towards the
   ring's background by f = (depth into the zone / width) dt r^-1.5; w relaxes to 0 */
void orc_damp(const pn_grid* g, const pn_params* P, void* rings, const pn_symdisk* val, double dt) {
  const int imax = g->imax, jmax = g->jmax, kmax = g->kmax;
  const double d_kill = 0.5 * P->sc_h;
  const double inner = P->xmin + d_kill, outer = P->xmax - d_kill;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < kmax; k++)
    for (int i = 0; i < imax; i++) {
      ring_t R = ring_at(rings, jmax, i + (long)imax * k);
      const double r = *R.xc;
      double f;
      if (r < inner) f = (inner - r) / d_kill;
      else if (r > outer) f = (r - outer) / d_kill;
      else continue;
      f *= dt * pow(r, -1.5);
      const pn_symdisk V = val[i + (long)imax * k];
      for (int j = 0; j < jmax; j++) {
        R.r[j] += f * (V.r - R.r[j]);
        R.p[j] += f * (V.p - R.p[j]);
        R.u[j] += f * (V.u - R.u[j]);
        R.v[j] += f * (V.v - R.v[j]);
        R.w[j] += f * (-R.w[j]);
      }
    }
}

/* the step loop of main.cu:343-387; with st->land, the endtime / tmovie landing of
   main.cu:352-366 (a cut step, then timem = 0 and dt = old_dt, main.cu:408-411), stopping
   after a step that landed with st->event = 1 (tmovie) or 2 (endtime) */
void orc_run(const pn_grid* g, const pn_params* P, void* rings, const void* lft, const void* rgh,
             const void* udr, const void* top, const pn_symdisk* val, int nsteps, pn_run_state* st) {
  st->event = 0;
  for (int s = 0; s < nsteps; s++) {
    double dt = orc_get_dt(g, P, rings, st->dt, st->FrRot);
    double old_dt = dt;
    int event = 0;
    if (st->land) {
      if (st->simtime + dt >= P->endtime) {
        old_dt = dt;
        dt = P->endtime - st->simtime;
        event = 2;
      }
      if (st->timem + dt >= P->tmovie) {
        old_dt = dt;
        dt = P->tmovie - st->timem;
        if (event == 0) event = 1;
      }
    }
    st->dt = dt;
    orc_index_shift(g, P, rings, val, dt, st->FrRot);
    if (P->plnt_flag > 0) orc_kick_drift_kick(P, &st->planet, &st->FrRot, 0.5 * dt, 1);
    orc_sweepx(g, P, rings, lft, rgh, dt, &st->planet, st->FrRot);
    orc_sweepy(g, P, rings, dt, &st->planet, st->FrRot);
    if (P->ndim == 3) orc_sweepz(g, P, rings, udr, top, dt, &st->planet, st->FrRot);
    if (P->fargo_flag > 0) orc_rotate(g, rings);
    if (P->kill_flag) orc_damp(g, P, rings, val, dt);
    st->simtime += dt;
    st->timem += dt;
    st->ncycle++;
    if (P->plnt_flag > 0) {
      orc_kick_drift_kick(P, &st->planet, &st->FrRot, dt, 0);
      orc_planet_forces(g, P, rings, &st->planet, st->FrRot, st->simtime);
      orc_kick_drift_kick(P, &st->planet, &st->FrRot, 0.5 * dt, 1);
    }
    if (event != 0) {
      st->timem = 0.0;
      st->dt = old_dt;
      st->event = event;
      return;
    }
  }
}
