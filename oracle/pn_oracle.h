/*
 * pn_oracle.h -- FP64 CPU restatement of PEnGUIn's hydro step.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as what results are compared
 * with (or the timed CPU baseline) -- never as the product path.
 *
 * The state is kept in the reference's own hydr_ring byte layout
 * (variable_types.h:67-97, runtime jmax) so that the oracle reads like the
 * reference operators and its bytes compare directly with pn_store_rings().
 */
#ifndef PN_ORACLE_H
#define PN_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "penguin.h"

#ifdef __cplusplus
extern "C" {
#endif

size_t orc_ring_bytes(int jmax);

/* init.cpp + disk_profile.cu + seeded noise; same contract as pn_init_disk */
int orc_init_disk(const pn_grid* g, const pn_params* p, uint64_t seed, double noise,
                  void* rings, void* lft, void* rgh, void* udr, void* top, pn_symdisk* val);

/* operators on hydr_ring arrays */
double orc_get_dt(const pn_grid* g, const pn_params* p, const void* rings, double dt_prev, double FrRot);
void orc_index_shift(const pn_grid* g, const pn_params* p, void* rings, const pn_symdisk* val,
                     double dt, double FrRot);
void orc_sweepx(const pn_grid* g, const pn_params* p, void* rings, const void* lft, const void* rgh,
                double dt, const pn_body* planet, double FrRot);
void orc_sweepy(const pn_grid* g, const pn_params* p, void* rings, double dt, const pn_body* planet,
                double FrRot);
void orc_sweepz(const pn_grid* g, const pn_params* p, void* rings, const void* udr, const void* top,
                double dt, const pn_body* planet, double FrRot);
void orc_rotate(const pn_grid* g, void* rings);
double orc_gravity(int geo, int comp, int bary, const pn_body* p, double rad, double azi, double pol, double dt);
void orc_damp(const pn_grid* g, const pn_params* p, void* rings, const pn_symdisk* val, double dt);
double orc_reduce(const pn_grid* g, const pn_params* p, const void* rings, const pn_body* planet,
                  double exclude, int mode);
void orc_totals(const pn_grid* g, const pn_params* p, const void* rings, double* mass, double* angmom);
/* the reader's dphi-weighted ring average of field 0..4 (r, p, u, v, w) per ring */
void orc_ring_means(const pn_grid* g, const void* rings, int field, double* out);
void orc_set_threads(int n);
int orc_get_threads(void);

/* planet.cu */
void orc_init_planet(const pn_params* p, pn_body* b, double* FrRot);
void orc_planet_forces(const pn_grid* g, const pn_params* p, const void* rings, pn_body* b,
                       double FrRot, double simtime);
void orc_kick_drift_kick(const pn_params* p, pn_body* b, double* FrRot, double dt, int kick);

/* main.cu loop body, nsteps times (no landing, no output) */
void orc_run_begin(const pn_grid* g, const pn_params* p, void* rings, const pn_symdisk* val,
                   pn_run_state* st);
void orc_run(const pn_grid* g, const pn_params* p, void* rings, const void* lft, const void* rgh,
             const void* udr, const void* top, const pn_symdisk* val, int nsteps, pn_run_state* st);

/* ---- building blocks, exposed for known-answer tests ------------------- */

/* paraset + parabola over a window of nmax cells with the index range
 * [nmin, nmax_p) exactly as states.cu (nmin 0) or remap.cu (nmin 3) calls
 * them.  Outputs al, da, a6 are written where the reference writes them and
 * left untouched elsewhere. */
void orc_parabola_window(int nmax, int nmin, int nmax_p, const double* dx, const double* val,
                         const double* flat, double* al, double* da, double* a6);
void orc_paraset_coeffs(int nmax, int nmin, int nmax_p, const double* dx, double* c /* 5*nmax */);
void orc_flatten_window(int nmax, const double* p, const double* u, double* flat);
void orc_riemann(int eos, double gam, double pl, double rl, double ul, double pr, double rr,
                 double ur, double* umid, double* pmid);
void orc_derivs(const double* xf, const double* u, int n, double* first, double* second);
double orc_force_rad_cyl(const pn_body* b, double rad, double azi, double z, double dt);
double orc_force_azi_cyl(const pn_body* b, double rad, double azi, double z, double dt);
double orc_potential(const pn_params* p, const pn_body* b, double rad, double azi, double z);
void orc_grid(int nzones, double xmin, double xmax, int lg, double* xa, double* xc, double* dx);

/* One directional PPMLR update of a window (ppmlr.cu + viscosity + EOS
 * closure), exposed so tests can drive single pencils: fields are in/out,
 * geometry per cell.  axis 0/1/2.  Returns nothing; cells [6, nmax-6) are
 * the updated interior. */
void orc_ppmlr_window(const pn_params* p, int axis, int nmax, double dt, double FrRot,
                      const pn_body* b, const double* rad, const double* azi, const double* pol,
                      const double* v_rot, const double* dvol0, double* xa0, double* dx0,
                      double* r, double* pr, double* u, double* v, double* w);

uint64_t orc_splitmix64(uint64_t x);
double orc_noise_u(uint64_t seed, uint64_t cell);

#ifdef __cplusplus
}
#endif

#endif
