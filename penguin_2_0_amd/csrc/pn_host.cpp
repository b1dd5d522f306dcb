// pn_host.cpp -- host-side pieces of the C ABI that touch no device: parameter defaults,
// the hydr_ring byte layout, synthetic disks (init.cpp / disk_profile.cu), the planet's
// kick-drift-kick (planet.cu) and the output files' names and headers (output.cu,
// global.cu).
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "penguin.h"
#include "pn_host.hpp"

namespace pn {

RingView ring_view(void* base, int jmax, long idx) {
  char* b = static_cast<char*>(base) + static_cast<size_t>(idx) * pn_ring_bytes(jmax);
  const size_t J = static_cast<size_t>(jmax);
  RingView v;
  v.r = reinterpret_cast<double*>(b);
  v.p = reinterpret_cast<double*>(b + 8 * J);
  v.u = reinterpret_cast<double*>(b + 16 * J);
  v.v = reinterpret_cast<double*>(b + 24 * J);
  v.w = reinterpret_cast<double*>(b + 32 * J);
  double* s0 = reinterpret_cast<double*>(b + 40 * J);
  v.x = s0;
  v.dx = s0 + 1;
  v.xvol = s0 + 2;
  v.xc = s0 + 3;
  v.y = reinterpret_cast<double*>(b + 40 * J + 32);
  v.dy = reinterpret_cast<double*>(b + 48 * J + 32);
  v.yvol = reinterpret_cast<double*>(b + 56 * J + 32);
  v.yc = reinterpret_cast<double*>(b + 64 * J + 32);
  double* s1 = reinterpret_cast<double*>(b + 72 * J + 32);
  v.z = s1;
  v.dz = s1 + 1;
  v.zvol = s1 + 2;
  v.zc = s1 + 3;
  v.i = reinterpret_cast<int32_t*>(b + 72 * J + 64);
  v.k = reinterpret_cast<int32_t*>(b + 72 * J + 68);
  v.rot_v = reinterpret_cast<double*>(b + 72 * J + 72);
  v.res_v = reinterpret_cast<double*>(b + 72 * J + 80);
  v.rot_j = reinterpret_cast<int32_t*>(b + 72 * J + 88);
  v.inc_j = reinterpret_cast<int32_t*>(b + 72 * J + 92);
  return v;
}

// ---- disk profile (disk_profile.cu:9-171) ----------------------------------------------

double Disk::cs2(double r) const { return P.sc_h * P.sc_h * std::pow(r, -P.p_beta); }
double Disk::h(double r) const { return std::sqrt(cs2(r) * r * r * r); }
double Disk::nu(double r) const { return P.vis_nu * std::pow(r, P.p_alpha); }

double Disk::rho(double r) const {
  double n = -P.p_alpha;
  if (P.eos != 0 && P.ndim == 3) n = -P.p_alpha * (2.0 / (P.gam + 1.0));
  return std::pow(r, n);
}

double Disk::pres(double r) const {
  if (P.eos == 0 || P.init_p_iso) return cs2(r) * rho(r);
  return cs2(r) * std::pow(rho(r), P.gam) / P.gam;
}

double Disk::dpdr(double r) const {
  const double p1 = pres(r + 1.27245e-8);
  const double p2 = pres(r - 1.27245e-8);
  return (p1 - p2) / 2.5449e-8;
}

double Disk::rho(double r, double z) const {
  if (P.eos == 0) {
    const double c2 = cs2(r);
    const double rr = std::sqrt(r * r + z * z);
    return rho(r) * std::exp((1.0 / rr - 1.0 / r) / c2);
  }
  const double gamm = P.gam - 1.0, gamfac2 = P.gam + 1.0;
  const double hh = std::sqrt(cs2(r) * r * r * r);
  const double z_0 = std::sqrt(2.0 * hh * hh * std::pow(r, -P.p_alpha * (2.0 / gamfac2) * gamm) / gamm);
  if (z < z_0)
    return std::pow(std::pow(rho(r), gamm) - (gamm / (r * cs2(r))) * (1.0 - r / std::sqrt(r * r + z * z)), 1.0 / gamm);
  return std::pow(10.0, std::log10(std::pow(std::pow(rho(r), gamm) -
                                                (gamm / (r * cs2(r))) * (1.0 - r / std::sqrt(r * r + z_0 * z_0)),
                                            1.0 / gamm)) -
                            6.0 * (z - z_0) / hh);
}

double Disk::pres(double r, double z) const {
  if (P.eos == 0 || P.init_p_iso) return cs2(r) * rho(r, z);
  return cs2(r) * std::pow(rho(r, z), P.gam) / P.gam;
}

double Disk::dpdr(double r, double z) const {
  const double p1 = pres(r + 1.27e-8, z);
  const double p2 = pres(r - 1.27e-8, z);
  return (p1 - p2) / 2.54e-8;
}

double Disk::dpds(double r, double pol) const {
  const double dr = 1.27e-7;
  const double sp = std::sin(pol), cp = std::cos(pol);
  const double r1 = (r + dr) * sp, z1 = (r + dr) * cp;
  const double r2 = (r - dr) * sp, z2 = (r - dr) * cp;
  return (pres(r1, z1) - pres(r2, z2)) / (2.0 * dr);
}

double Disk::viscous_vr(double r) const { return (P.visc_flag == 1) ? -1.5 * nu(r) / r : 0.0; }

// ---- grid (grid.cu:67-248) -------------------------------------------------------------

void make_grid(int nz, double lo, double hi, int lg, double* xa, double* xc, double* dx) {
  if (lg == 0) {
    const double f = (hi - lo) / static_cast<double>(nz);
    for (int n = 0; n < nz; ++n) {
      xa[n] = lo + static_cast<double>(n) * f;
      dx[n] = f;
      xc[n] = xa[n] + 0.5 * dx[n];
    }
    return;
  }
  const double f = (std::log(hi) - std::log(lo)) / static_cast<double>(nz);
  xa[0] = lo;
  for (int n = 1; n < nz; ++n) {
    xa[n] = lo * std::exp(f * static_cast<double>(n));
    dx[n - 1] = xa[n] - xa[n - 1];
    xc[n - 1] = xa[n - 1] + 0.5 * dx[n - 1];
  }
  dx[nz - 1] = hi - xa[nz - 1];
  xc[nz - 1] = xa[nz - 1] + 0.5 * dx[nz - 1];
}

double cell_vol(const pn_params& P, double xa, double dx, int axis) {
  if (axis == 0) {
    if (P.ngeomx == 1) return dx * (xa + 0.5 * dx);
    if (P.ngeomx == 2) return dx * (xa * (xa + dx) + dx * dx * 0.3333333333333333333333333);
    return dx;
  }
  if (axis == 1) return dx;
  if (P.ndim == 2) return 1.0;
  if (P.ngeomz == 5) return std::cos(xa) - std::cos(xa + dx);
  return dx;
}

uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double noise_u(uint64_t seed, uint64_t cell) {
  return static_cast<double>(splitmix64(seed ^ cell) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

}  // namespace pn

using namespace pn;

extern "C" {

void pn_params_default(pn_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->ndim = 2;
  p->eos = 0;
  p->flat_flag = 1;
  p->plnt_flag = 2;
  p->frrot_flag = 1;
  p->visc_flag = 1;
  p->fargo_flag = 1;
  p->ngeomx = 1;
  p->ngeomy = 3;
  p->ngeomz = 0;
  p->grspx = 1;
  p->grspy = 0;
  p->grspz = 0;
  p->init_p_iso = 0;
  p->p_alpha = 1.5;
  p->p_beta = 0.5;
  p->ss_alpha = 0.0001;
  p->sc_h = 0.05;
  p->vis_nu = p->ss_alpha * p->sc_h * p->sc_h;
  p->M_p = 1.0 * 0.001;
  p->R_p = 1.0;
  p->t_growth = 1000.0 * 6.283185307179586476925286766559;
  p->rs_fac = 0.5;
  p->gam = 1.0;
  p->courant = 0.5;
  p->endtime = 10000.0 * 6.283185307179586476925286766559;
  p->tmovie = 200.0 * 6.283185307179586476925286766559;
  p->xmin = 0.2;
  p->xmax = 3.0;
  p->ymin = 0.0;
  p->ymax = 6.283185307179586476925286766559;
  p->zmin = 1.570796326794896619231321691639751442 - 0.065;
  p->zmax = 1.570796326794896619231321691639751442;
  p->nlft = 3;
  p->nrgh = 3;
  p->kill_flag = 0;
  p->bary_flag = 0;
}

size_t pn_ring_bytes(int32_t jmax) { return static_cast<size_t>(72) * static_cast<size_t>(jmax) + 96; }

// update_grid (read/p_init.cpp:108-132): lab column j of ring (i, k) holds data index
// j - rot_j (mod jmax); outputs are [i][j][k]
int pn_lab_fields(const void* rings, size_t bytes, int32_t imax, int32_t jmax, int32_t kmax, double* r, double* p,
                  double* u, double* v, double* w) {
  if (!rings || imax < 1 || jmax < 1 || kmax < 1) return PN_ERR_ARG;
  const size_t rb = pn_ring_bytes(jmax);
  const long nring = static_cast<long>(imax) * kmax;
  if (bytes != rb * static_cast<size_t>(nring)) return PN_ERR_ARG;
  std::vector<char> seen(static_cast<size_t>(nring), 0);
  for (long n = 0; n < nring; ++n) {
    RingView R = ring_view(const_cast<void*>(rings), jmax, n);
    const int i = *R.i, k = *R.k;
    if (i < 0 || i >= imax || k < 0 || k >= kmax || seen[static_cast<size_t>(i) + static_cast<size_t>(imax) * k])
      return PN_ERR_ARG;
    seen[static_cast<size_t>(i) + static_cast<size_t>(imax) * k] = 1;
    const long rot = *R.rot_j % jmax;
    const double* src[5] = {R.r, R.p, R.u, R.v, R.w};
    double* dst[5] = {r, p, u, v, w};
    for (int f = 0; f < 5; ++f) {
      if (!dst[f]) continue;
      for (int j = 0; j < jmax; ++j) {
        long j0 = j - rot;
        if (j0 < 0) j0 += jmax;
        if (j0 >= jmax) j0 -= jmax;
        dst[f][(static_cast<size_t>(i) * jmax + j) * kmax + k] = src[f][j0];
      }
    }
  }
  return PN_OK;
}

// init_grid / init_den / init_speed / init_cells / init_bound (init.cpp:3-543)
int pn_init_disk(const pn_grid* g, const pn_params* p, uint64_t seed, double noise, void* rings, void* lft,
                 void* rgh, void* udr, void* top, pn_symdisk* val) {
  if (!g || !p || !rings || !lft || !rgh || !val) return PN_ERR_ARG;
  const int I = g->imax, J = g->jmax, K = g->kmax;
  if (I < 1 || J < 2 * PN_NPAD || K < 1) return PN_ERR_ARG;
  if ((p->ndim == 2) != (K == 1)) return PN_ERR_ARG;
  if (p->ndim == 3 && (!udr || !top)) return PN_ERR_ARG;
  const pn_params& P = *p;
  Disk D{P};
  std::vector<double> xa(I), xc(I), dx(I), ya(J), yc(J), dy(J), za(K), zc(K), dz(K);
  make_grid(I, P.xmin, P.xmax, P.grspx, xa.data(), xc.data(), dx.data());
  make_grid(J, P.ymin, P.ymax, P.grspy, ya.data(), yc.data(), dy.data());
  make_grid(K, P.zmin, P.zmax, P.grspz, za.data(), zc.data(), dz.data());

  // background per ring: density, pressure, v_phi (pressure-corrected Keplerian), v_r
  for (int i = 0; i < I; ++i)
    for (int k = 0; k < K; ++k) {
      pn_symdisk& s = val[i + I * k];
      double r = xc[i], z = zc[k], vk2, dP;
      if (P.ngeomz == 5) {
        r = xc[i] * std::sin(zc[k]);
        z = xc[i] * std::cos(zc[k]);
      }
      if (P.ndim == 3) {
        s.r = D.rho(r, z);
        s.p = D.pres(r, z);
      } else {
        s.r = D.rho(r);
        s.p = D.pres(r);
      }
      if (P.ndim == 2) {
        r = xc[i];
        vk2 = 1.0 / r;
        dP = D.dpdr(xc[i]);
      } else if (P.ngeomz == 5) {
        vk2 = 1.0 / xc[i];
        dP = D.dpds(xc[i], zc[k]);
      } else {
        dP = D.dpdr(xc[i], zc[k]);
        vk2 = r * r * std::pow(r * r + z * z, -1.5);
      }
      vk2 += xc[i] * dP / s.r;
      s.v = std::sqrt(vk2);
      s.u = D.viscous_vr(r);
      s.w = 0.0;
    }

  const size_t rb = pn_ring_bytes(J);
  std::memset(rings, 0, rb * static_cast<size_t>(I) * K);
  for (int k = 0; k < K; ++k)
    for (int i = 0; i < I; ++i) {
      const long m = i + static_cast<long>(I) * k;
      RingView R = ring_view(rings, J, m);
      const pn_symdisk& s = val[i + I * k];
      *R.k = k;
      *R.z = za[k];
      *R.dz = dz[k];
      *R.zc = *R.z + 0.5 * *R.dz;
      *R.i = i;
      *R.x = xa[i];
      *R.dx = dx[i];
      *R.xc = *R.x + 0.5 * *R.dx;
      const double rad_cyl = (P.ngeomz == 5) ? *R.xc * std::sin(*R.zc) : *R.xc;
      for (int j = 0; j < J; ++j) {
        R.y[j] = ya[j];
        R.dy[j] = dy[j];
        R.yc[j] = R.y[j] + 0.5 * R.dy[j];
        R.yvol[j] = cell_vol(P, ya[j], dy[j], 1);
        double rho = s.r;
        double pr = s.p;
        if (noise != 0.0) {
          const uint64_t cell = static_cast<uint64_t>(j) + static_cast<uint64_t>(J) * static_cast<uint64_t>(m);
          rho = rho * (1.0 + noise * noise_u(seed, cell));
          pr = (P.eos == 0 || P.init_p_iso) ? rho * D.cs2(rad_cyl) : D.cs2(rad_cyl) * std::pow(rho, P.gam) / P.gam;
        }
        R.r[j] = rho;
        R.p[j] = pr;
        R.u[j] = s.u;
        R.v[j] = s.v;
        R.w[j] = 0.0;
      }
      *R.xvol = cell_vol(P, *R.x, *R.dx, 0);
      *R.zvol = cell_vol(P, *R.z, *R.dz, 2);
    }

  // fixed ghost rings: 6 per radial side per theta layer, 6 per theta side per ring
  auto ghost = [&](RingView R, double gx, double gdx, double gz, double gdz, double gzc, bool theta_side) {
    *R.x = gx;
    *R.dx = gdx;
    *R.xc = theta_side ? gx + 0.5 * gdx : *R.x + 0.5 * *R.dx;
    *R.xvol = cell_vol(P, gx, gdx, 0);
    *R.z = gz;
    *R.dz = gdz;
    *R.zc = gzc;
    *R.zvol = cell_vol(P, gz, gdz, 2);
    const double rr = *R.xc;
    double r = rr, z = 0.0, rho, pr, dP, vk;
    if (P.ndim == 2) {
      rho = D.rho(r);
      pr = D.pres(r);
      dP = D.dpdr(r);
      vk = 1.0 / r;
    } else if (P.ngeomz == 5) {
      r = rr * std::sin(gzc);
      z = rr * std::cos(gzc);
      if (theta_side) r = std::sqrt(rr * rr - z * z);
      rho = D.rho(r, z);
      pr = D.pres(r, z);
      dP = D.dpds(rr, gzc);
      vk = 1.0 / rr;
    } else {
      z = gzc;
      rho = D.rho(r, z);
      pr = D.pres(r, z);
      dP = D.dpdr(r, z);
      vk = theta_side ? 1.0 / rr : r * r * std::pow(r * r + z * z, -1.5);
    }
    vk += rr * dP / rho;
    vk = std::sqrt(vk);
    for (int j = 0; j < J; ++j) {
      R.r[j] = rho;
      R.p[j] = pr;
      R.u[j] = D.viscous_vr(r);
      R.v[j] = vk;
      R.w[j] = 0.0;
      R.y[j] = ya[j];
      R.dy[j] = dy[j];
      R.yc[j] = yc[j];
      R.yvol[j] = cell_vol(P, ya[j], dy[j], 1);
    }
    *R.rot_v = (P.fargo_flag == 1) ? vk - 1.0 * *R.xc : 0.0;   // init.cpp:323 (unit frame rate)
    *R.rot_j = 0;
    *R.res_v = 0.0;
    *R.inc_j = 0;
    *R.i = 0;
    *R.k = 0;
  };
  std::memset(lft, 0, rb * PN_NPAD * static_cast<size_t>(K));
  std::memset(rgh, 0, rb * PN_NPAD * static_cast<size_t>(K));
  for (int i = 0; i < PN_NPAD; ++i)
    for (int k = 0; k < K; ++k) {
      const int ib = i + PN_NPAD * k;
      ghost(ring_view(lft, J, ib), xa[0] - dx[0] * static_cast<double>(PN_NPAD - i), dx[0], za[k], dz[k], zc[k], false);
      ghost(ring_view(rgh, J, ib), xa[I - 1] + dx[I - 1] * static_cast<double>(i + 1), dx[I - 1], za[k], dz[k], zc[k],
            false);
    }
  if (P.ndim == 3) {
    std::memset(udr, 0, rb * PN_NPAD * static_cast<size_t>(I));
    std::memset(top, 0, rb * PN_NPAD * static_cast<size_t>(I));
    for (int k = 0; k < PN_NPAD; ++k)
      for (int i = 0; i < I; ++i) {
        const int ib = k + PN_NPAD * i;
        double z0 = za[0] - dz[0] * static_cast<double>(PN_NPAD - k);
        double z1 = za[K - 1] + dz[K - 1] * static_cast<double>(k + 1);
        ghost(ring_view(udr, J, ib), xa[i], dx[i], z0, dz[0], z0 + 0.5 * dz[0], true);
        ghost(ring_view(top, J, ib), xa[i], dx[i], z1, dz[K - 1], z1 + 0.5 * dz[K - 1], true);
      }
  }
  return PN_OK;
}

// ---- planet (planet.cu:28-93) -----------------------------------------------------------

static double set_M_p(const pn_params* P, double t) {
  if (t < P->t_growth) return std::fmax(P->M_p * std::sin(3.1415926535897932384626433832795 * t / (2.0 * P->t_growth)), 0.01 * 0.000003);
  return P->M_p;
}

void pn_init_planet(const pn_params* p, pn_body* b, double* FrRot) {
  if (!p || !b || !FrRot) return;
  Disk D{*p};
  std::memset(b, 0, sizeof *b);
  b->m = set_M_p(p, 0.0);
  b->x = p->R_p;
  b->y = 3.1415926535897932384626433832795;
  b->vx = 0.0;
  b->vy = std::pow(b->x, -1.5);
  *FrRot = (p->frrot_flag == 1) ? b->vy : 0.0;
  b->vy -= *FrRot;
  b->rs = p->rs_fac * D.h(b->x);
}

void pn_kick_drift_kick(const pn_params* p, pn_body* b, double* FrRot, double dt, int kick) {
  if (!p || !b || !FrRot) return;
  if (!kick) {
    Disk D{*p};
    const double old_rad = b->x;
    b->x += dt * b->vx;
    b->y += dt * b->vy;
    b->vy = ((b->vy + *FrRot) * old_rad * old_rad / b->x / b->x) - *FrRot;
    if (b->y < 0.0) b->y += 6.283185307179586476925286766559;
    if (b->y >= 6.283185307179586476925286766559) b->y -= 6.283185307179586476925286766559;
    b->rs = p->rs_fac * D.h(b->x);
  } else {
    b->vx += dt * b->fx;
    b->vy += *FrRot + dt * b->fy / b->x / b->x;
    *FrRot = (p->frrot_flag == 1) ? b->vy : 0.0;
    b->vy -= *FrRot;
  }
}

// ---- output names and headers (output.cu:57-120, global.cu:20-38, main.cu:113-124) ------

// copies s with its terminating NUL into buf[n]; PN_ERR_ARG when it does not fit
static int put_string(const std::string& s, char* buf, size_t n) {
  if (!buf || s.size() + 1 > n) return PN_ERR_ARG;
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return PN_OK;
}

// the reference's int_to_string takes an int: a double argument is truncated towards zero
static std::string istr(double v) { return std::to_string(static_cast<int>(v)); }

int pn_label(const pn_grid* g, const pn_params* p, char* buf, size_t n) {
  if (!g || !p) return PN_ERR_ARG;
  std::string l = istr(g->imax) + "x" + istr(g->jmax) + "x" + istr(g->kmax) + "_c" + istr(std::round(p->sc_h / 0.001));
  // (the opacity tags of opac_flag have no counterpart: that module is not part of this build)
  if (p->plnt_flag == 2) l += "_pm" + istr(std::round(p->M_p / 0.000003)) + "_s" + istr(std::round(100.0 * p->rs_fac));
  else if (p->plnt_flag == 1) l += "_p" + istr(std::round(p->M_p / 0.000003)) + "_s" + istr(std::round(100.0 * p->rs_fac));
  l += "_i" + istr(p->xmin * 100.0) + "_o" + istr(p->xmax * 100.0);
  if (p->visc_flag == 1) l += "_a" + istr(std::round(p->ss_alpha * 10000.0));
  if (p->eos == 1) {
    l += "_g" + istr(std::round(p->gam * 10.0));
  } else if (p->eos == 0) {
    if (p->p_alpha + 0.5 * p->p_beta - 1.5 == 1.5) l += "_flatvor";
    if (p->p_beta == 0.0) l += "_iso";
    else l += "_b" + istr(p->p_beta * 10.0);
  } else if (p->eos == 2) {
    l += "_ad";
  } else {
    l += "_ad3";   // EOS 3 has no tag in the reference
  }
  if (p->frrot_flag == 1) l += "_rot";
  if (p->fargo_flag == 1) l += "_fargo";
  return put_string(l, buf, n);
}

int pn_frame_name(int num, char* buf, size_t n) {
  // five digits, zero padded; the reference returns an empty name past 99999 (and for a
  // negative number), which would overwrite one file: refused here
  if (num < 0 || num > 99999) return PN_ERR_ARG;
  std::string c = std::to_string(num);
  return put_string(std::string(5 - c.size(), '0') + c, buf, n);
}

int pn_write_para(const pn_grid* g, const pn_params* p, const char* mainp, const char* label, const char* path) {
  if (!g || !p || !mainp || !label || !path) return PN_ERR_ARG;
  std::ofstream f(path);
  if (!f) return PN_ERR_IO;
  // one GPU holds the whole disk: nDev = 1; default stream formatting, as ofstream << gives
  f << g->imax << "\n" << g->jmax << "\n" << g->kmax << "\n" << 1 << "\n" << p->gam << "\n" << p->sc_h << "\n"
    << p->tmovie << "\n" << mainp << "\n" << label << "\n" << p->grspx << "\n" << p->grspy << "\n" << p->grspz << "\n"
    << p->xmin << " " << p->xmax << "\n" << p->ymin << " " << p->ymax << "\n" << p->zmin << " " << p->zmax << "\n";
  return f.good() ? PN_OK : PN_ERR_IO;
}

int pn_write_initial(const pn_grid* g, const pn_params* p, const char* path) {
  if (!g || !p || !path || g->imax < 1) return PN_ERR_ARG;
  std::vector<double> xa(g->imax), xc(g->imax), dx(g->imax);
  pn::make_grid(g->imax, p->xmin, p->xmax, p->grspx, xa.data(), xc.data(), dx.data());
  std::ofstream f(path);
  if (!f) return PN_ERR_IO;
  for (int i = 0; i < g->imax; ++i) f << xa[i] << " " << dx[i] << "\n";
  return f.good() ? PN_OK : PN_ERR_IO;
}

}  // extern "C"
