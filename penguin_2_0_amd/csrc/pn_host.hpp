// pn_host.hpp -- host helpers: hydr_ring views, disk profile, grid.
#pragma once

#include <stdint.h>

#include "penguin.h"

namespace pn {

// pointers into one hydr_ring record (variable_types.h:67-97) for a runtime jmax
struct RingView {
  double *r, *p, *u, *v, *w;
  double *x, *dx, *xvol, *xc;
  double *y, *dy, *yvol, *yc;
  double *z, *dz, *zvol, *zc;
  int32_t *i, *k;
  double *rot_v, *res_v;
  int32_t *rot_j, *inc_j;
};

RingView ring_view(void* base, int jmax, long idx);

// analytic disk of disk_profile.cu:9-171 for one parameter set
struct Disk {
  const pn_params& P;
  double cs2(double r) const;
  double h(double r) const;
  double nu(double r) const;
  double rho(double r) const;
  double pres(double r) const;
  double dpdr(double r) const;
  double rho(double r, double z) const;
  double pres(double r, double z) const;
  double dpdr(double r, double z) const;
  double dpds(double r, double pol) const;
  double viscous_vr(double r) const;
};

void make_grid(int nz, double lo, double hi, int lg, double* xa, double* xc, double* dx);
double cell_vol(const pn_params& P, double xa, double dx, int axis);
uint64_t splitmix64(uint64_t x);
double noise_u(uint64_t seed, uint64_t cell);

}  // namespace pn
