/*
 * penguin.h -- C ABI of the MI355X-native PEnGUIn hydro step.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (PN_OK == 0) and never exits the process.  Detail for a non-zero
 * status is available from pn_last_error(ctx).  Each call completes (its
 * device work has drained) before it returns, as the reference operators do.
 * A context is not thread-safe; separate contexts on separate threads are.
 *
 * Reference interface replaced (paths relative to the reference checkout):
 *   types   hydr_ring / body / SymDisk      include/variable_types.h:26-97
 *   params  compile-time flags + constants  include/global.h:16-96, variable_types.h:11-20
 *   ops     get_dt                          include/timestep.h:4        (src/timestep/timestep.cu:112-149)
 *           Index_Shift                     include/orbital_advection.h:4 (src/orbital_advection/orbital_advection.cu:12-69)
 *           bundle_sweep2D                  include/ppm.h:4             (src/ppm/ppm.cu:358-386)
 *           GPU_output_reduction            include/output.h:32         (src/output/device_func.cu:49-198)
 *           save_cells / copy_bfile         include/output.h:26         (src/output/output.cu:48-55, main.cu:37-42)
 *           init_planet / planet_forces /
 *           kick_drift_kick                 include/planet.h:7-9        (src/planet/planet.cu:36-93)
 *   loop    the step loop of kernel()       main.cu:343-441
 */
#ifndef PENGUIN_H
#define PENGUIN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PN_NPAD 6          /* ghost rings / halo cells per side, variable_types.h:13 */

/* status codes */
enum {
  PN_OK = 0,
  PN_ERR_ARG = 1,          /* bad argument: null pointer, size mismatch, out-of-range flag */
  PN_ERR_HIP = 2,          /* a HIP runtime call failed (message carries hipGetErrorString) */
  PN_ERR_LAYOUT = 3,       /* hydr_ring bytes inconsistent (ring index, phi geometry) */
  PN_ERR_STATE = 4,        /* call out of order (e.g. sweep before load) */
  PN_ERR_UNSUPPORTED = 5,  /* flag combination this build does not implement */
  PN_ERR_NUMERIC = 6,      /* non-finite dt or state detected */
  PN_ERR_IO = 7            /* snapshot file missing, short or unwritable */
};

/* arithmetic precision of a context */
enum {
  PN_FP64 = 0,             /* plain FP64: the drop-in path */
  PN_FP32 = 1              /* FP32 bulk with the FP64 steps kept where stability needs them */
};

/* body, variable_types.h:26-37 (same field order) */
typedef struct pn_body {
  double m, x, y, vx, vy, fx, fy, rs;
} pn_body;

/* SymDisk, variable_types.h:58-65: the analytic background per ring (FARGO speed) */
typedef struct pn_symdisk {
  double r, p, u, v, w;
} pn_symdisk;

/* grid sizes, variable_types.h:18-20; kmax == 1 means a 2-D r-phi disk */
typedef struct pn_grid {
  int32_t imax, jmax, kmax;
} pn_grid;

/* The compile-time flags and constants of global.h as one runtime record.
 * pn_params_default() fills the values global.h ships with. */
typedef struct pn_params {
  /* flags, global.h:16-26 */
  int32_t ndim;            /* 2 or 3 (variable_types.h:11) */
  int32_t eos;             /* 0 isothermal, 1 isentropic, 2 adiabatic (global.h:18); 3 adiabatic
                              with the energy-consistent Lagrangian step that evolve.cu:146 leaves
                              commented out (2-D; not in the reference's live code).  PN_FP32
                              contexts take 0 and 3 */
  int32_t flat_flag;       /* 1: PPM flattening at shocks (global.h:16) */
  int32_t plnt_flag;       /* 0 none, 1 fixed, 2 moving planet (global.h:20) */
  int32_t frrot_flag;      /* 1: frame co-rotates with the planet (global.h:21) */
  int32_t visc_flag;       /* 1: explicit viscosity (global.h:23) */
  int32_t fargo_flag;      /* 0 off, 1 background orbital advection, 2 frame (global.h:24) */
  int32_t ngeomx, ngeomy, ngeomz;  /* geometry codes (global.h:83-85) */
  int32_t grspx, grspy, grspz;     /* grid spacing codes (global.h:87-89) */
  int32_t init_p_iso;      /* 1: initial P = c_s^2 Sigma for every EOS (else disk_profile.cu:54-61) */
  /* disk, global.h:32-37 */
  double p_alpha, p_beta, ss_alpha, sc_h, vis_nu;
  /* planet, global.h:43-46 */
  double M_p, R_p, t_growth, rs_fac;
  /* hydro, global.h:52-57 */
  double gam, courant;
  double endtime, tmovie;
  /* grid bounds, global.h:76-81 */
  double xmin, xmax, ymin, ymax, zmin, zmax;
  /* radial boundaries, global.h:59-60, with the meanings set_boundx gives the codes
   * (boundary.cu:86-206; the comment at global.h:59 lists them differently):
   * 0 reflect (ghost t cells outside the edge cell take the cell t inside it, v_r negated),
   * 1 outflow (ghosts copy the edge cell), 2 periodic in radius, 3 fixed ghost rings (the
   * reference's only live choice, the default), 4 outflow of r, p, u (left) or r, p (right)
   * with the fixed ghosts' other fields.  Both paths; on the FP32 path a ghost ring that a
   * code copies into carries its source ring's backgrounds, so the copies are of deviations. */
  int32_t nlft, nrgh;
  /* 1: damping zones of width sc_h/2 at xmin and xmax relaxing every field towards the
   * background at rate (distance into the zone / width) r^-1.5 per unit time
   * (kill_wave.cu:3-43; kill_flag, global.h:22).  Both paths; the FP32 path scales the
   * deviations from the ring backgrounds. */
  int32_t kill_flag;
  /* 1: barycentric frame (bary_flag, global.h:17): star and planet orbit their barycentre,
   * the star's pull is computed from its offset position and the indirect term goes
   * (forces.cu:28-182).  Both paths; on the FP32 path the star's offset pull is a deviation
   * force beside the planet's, in kernels of their own (the disks of a batch share the frame). */
  int32_t bary_flag;
} pn_params;

/* State a step loop carries between steps (main.cu:270-275, 325-327). */
typedef struct pn_run_state {
  double dt;               /* previous dt (endtime before the first step, main.cu:270) */
  double simtime;
  double FrRot;            /* frame rotation rate */
  pn_body planet;
  int64_t ncycle;
  double timem;            /* time since the last picture (main.cu:273) */
  int32_t land;            /* in: 1 applies the endtime / tmovie landing of main.cu:352-366 */
  int32_t event;           /* out: 0, or the landing the last step made: 1 tmovie (a picture
                              is due), 2 endtime (the run is over) */
} pn_run_state;

typedef struct pn_ctx pn_ctx;

/* ---- parameters and layout ------------------------------------------- */

/* Fill p with the values of global.h / variable_types.h. */
void pn_params_default(pn_params* p);

/* sizeof(hydr_ring) for a given jmax: 72*jmax + 96 (variable_types.h:67-97). */
size_t pn_ring_bytes(int32_t jmax);

/* The snapshot reader's un-rotation (update_grid, read/p_init.cpp:108-132): from the
 * imax*kmax hydr_ring records in `rings` (any ring order; each record names its i and k)
 * fill lab-frame arrays indexed [i][j][k] -- element (i*jmax + j)*kmax + k -- with lab
 * column j taken from data index (j - rot_j) mod jmax.  Any output may be NULL.  Host only.
 * PN_ERR_ARG on a size mismatch or a ring whose (i, k) is out of range or repeated. */
int pn_lab_fields(const void* rings, size_t bytes, int32_t imax, int32_t jmax, int32_t kmax,
                  double* r, double* p, double* u, double* v, double* w);

/* ---- host-side synthetic disk (init.cpp:3-543, disk_profile.cu:9-171) --
 * Writes imax*kmax interior rings (ring order i + imax*k, init.cpp:241),
 * PN_NPAD*kmax fixed ghost rings on each radial side (lft, rgh), and, in 3-D,
 * PN_NPAD*imax fixed ghost rings below/above in theta (udr, top; may be NULL
 * in 2-D).  val receives imax*kmax background records.  Surface density gets
 * seeded relative noise Sigma *= 1 + noise*U, U = splitmix64(seed ^ cell)
 * mapped to [-1, 1), cell = j + jmax*(i + imax*k).  Before any step the
 * caller primes FARGO with pn_index_shift(ctx, endtime, FrRot) as main.cu:330
 * does.  Returns PN_OK or PN_ERR_ARG. */
int pn_init_disk(const pn_grid* g, const pn_params* p, uint64_t seed, double noise,
                 void* rings, void* lft, void* rgh, void* udr, void* top,
                 pn_symdisk* val);

/* ---- context ----------------------------------------------------------- */

/* Allocate a context on the current HIP device.  precision: PN_FP64 or
 * PN_FP32.  The library owns all device memory; the caller owns host
 * buffers. */
int pn_create(const pn_grid* g, const pn_params* p, int precision, pn_ctx** out);
int pn_destroy(pn_ctx* ctx);
const char* pn_last_error(const pn_ctx* ctx);

/* Upload state in hydr_ring byte layout.  rings: imax*kmax records;
 * lft/rgh: PN_NPAD*kmax records each; udr/top: PN_NPAD*imax records each
 * (3-D only, else NULL with udr_bytes 0); val: imax*kmax records.  All rings
 * must carry the same phi geometry (y, dy, yc, yvol) -- the reference always
 * builds them that way (init.cpp:201-206) -- else PN_ERR_LAYOUT. */
int pn_load_rings(pn_ctx* ctx, const void* rings, size_t rings_bytes,
                  const void* lft, const void* rgh, size_t lr_bytes,
                  const void* udr, const void* top, size_t ut_bytes,
                  const pn_symdisk* val, size_t nval);

/* Download interior rings in hydr_ring byte layout (the bytes save_cells
 * writes, output.cu:48-55). */
int pn_store_rings(pn_ctx* ctx, void* rings, size_t rings_bytes);

/* ---- step operators ----------------------------------------------------- */

/* get_dt (timestep.cu:112-149): *dt = min(1.1*dt_prev, courant / max over
 * cells of the signal speed per cell width). */
int pn_get_dt(pn_ctx* ctx, double dt_prev, double FrRot, double* dt);

/* Index_Shift (orbital_advection.cu:12-43): per ring, split the previous
 * orbital shift into whole cells (inc_j) and a residual speed (res_v), then
 * refresh rot_v. */
int pn_index_shift(pn_ctx* ctx, double dt, double FrRot);

/* bundle_sweep2D (ppm.cu:358-386): radial sweep, azimuthal sweep, integer
 * phi roll.  For a 3-D context it also runs the theta sweep between the
 * azimuthal sweep and the roll. */
int pn_sweep2d(pn_ctx* ctx, double dt, const pn_body* planet, double FrRot);

/* One directional sweep (axis 0 = r, 1 = phi, 2 = theta) without the roll.  No
 * counterpart in the reference, which only exposes the bundle; it lets callers and tests
 * drive a single direction (ppm.cu:26-160 or 162-307). */
int pn_sweep_axis(pn_ctx* ctx, int axis, double dt, const pn_body* planet, double FrRot);

/* GPU_output_reduction (device_func.cu:49-198): mode 0 = total mass,
 * 8 / 9 = background-subtracted radial / azimuthal force on the planet
 * (2-D).  exclude is in units of sc_h as in the reference. */
int pn_reduce(pn_ctx* ctx, const pn_body* planet, double exclude, int mode, double* out);

/* Azimuthal mean of one field of every interior ring, into out[i + imax*k] (n = imax*kmax):
 * sum_j f[j] dphi[j0] / sum_j dphi[j0] over the ring's data columns j at lab columns
 * j0 = j + rot_j -- the dphi-weighted ring average of the reader (write_output_file,
 * read/p_output.cpp:112-127, which divides by 2 pi, the sum of dphi over a full ring).
 * field: 0 Sigma (r), 1 P (p), 2 v_r (u), 3 v_phi (v), 4 w (v_theta in 3-D).  A wavefront-
 * shuffle reduction on the device, FP64 sums in a fixed order; on an FP32 context the
 * deviations from the ring's background are summed and the background added back.  A
 * diagnostic beside the step: the FARGO background stays the initial analytic v_phi(r)
 * (orbital_advection.cu:21-22), as in the reference. */
int pn_ring_means(pn_ctx* ctx, int field, double* out, size_t n);

/* ---- planet (host, planet.cu:28-93) ------------------------------------ */
void pn_init_planet(const pn_params* p, pn_body* planet, double* FrRot);
/* mass ramp + forces; for plnt_flag == 2 it runs two pn_reduce calls. */
int pn_planet_forces(pn_ctx* ctx, pn_body* planet, double FrRot, double simtime);
void pn_kick_drift_kick(const pn_params* p, pn_body* planet, double* FrRot, double dt, int kick);

/* ---- the step loop ------------------------------------------------------- */

/* Prime a run: st->dt = endtime, planet initialised, FARGO primed
 * (main.cu:270, 325-330). */
int pn_run_begin(pn_ctx* ctx, pn_run_state* st);
/* Prime a run resumed from picture startat (restart, main.cu:244-330): st->dt = endtime,
 * simtime = tmovie * startat, timem = 0, the planet initialised afresh with its mass ramp
 * at that simtime (the snapshots hold no planet), and no priming Index_Shift: the rings
 * keep the FARGO state they were saved with (main.cu:330).  Load the picture with
 * pn_load_snapshot first.  PN_ERR_ARG for startat < 0. */
int pn_run_restart(pn_ctx* ctx, int startat, pn_run_state* st);

/* Advance nsteps steps of main.cu:343-387 (no output), keeping dt, the planet and FrRot on
 * the device between steps.  The step is launched as a captured HIP graph when graph != 0.
 * With st->land set, each step applies the landing of main.cu:352-366 on the device (a step
 * that would pass endtime or the next picture time, timem + dt >= tmovie, is cut to land on
 * it; after it timem restarts and dt returns to the CFL value), and pn_run returns right
 * after such a step with st->event = 1 (picture due) or 2 (endtime reached) and st->ncycle
 * counting the steps done.  To stop there without a device-side halt, the host enqueues at
 * most as many steps as cannot reach a landing (dt grows by at most 1.1 per step) and reads
 * the scalars between those batches.  PN_ERR_ARG when st->land is set and no time is left
 * (timem >= tmovie or simtime >= endtime): that step would be cut to dt <= 0. */
int pn_run(pn_ctx* ctx, int nsteps, int graph, pn_run_state* st);

/* Time the kernels of one step with HIP events on the context's stream: each of
 * {CFL, Index_Shift, r sweep, phi sweep, theta sweep, roll} is launched nrep times untimed
 * (clocks back to speed after host-side idle), then nrep times back to back between two
 * events, and its mean duration in milliseconds written to ms[0..5]
 * (0 where a kernel does not run).  No counterpart in the reference (which only prints
 * seconds per step, main.cu:430-437).  The state advances while it is timed. */
int pn_time_kernels(pn_ctx* ctx, int nrep, double* ms6);

/* Autotune the FP32 sweeps: time every workgroup-shape variant of the radial sweep (in
 * 3-D: radial + theta sweep, which share it) and of the azimuthal sweep on the current
 * state (nrep launches each, state left unchanged) and keep the fastest.  *var_x / *var_y
 * receive the choice (may be NULL).  FP64 contexts have one variant (0).  pn_run tunes an
 * FP32 context on its first call.  No counterpart in the reference. */
int pn_tune(pn_ctx* ctx, int nrep, int* var_x, int* var_y);
/* Pin the FP32 sweep variants instead of tuning.  var_x 0..3, radial / theta sweeps as
 * (cells per lane, waves per workgroup, lab columns per wave): (2,4,2), (2,8,1), (2,16,1),
 * (2,4,4).  var_y 0..3, azimuthal sweep as (cells per lane, waves per workgroup): (2,2),
 * (2,4), (4,2), (4,4); a viscous disk takes var_y 0 or 1 only (PN_ERR_ARG otherwise), since
 * the viscosity after the remap reads halo cells at window edges and so depends on the
 * reference's 128-cell windows. */
int pn_set_variants(pn_ctx* ctx, int var_x, int var_y);
/* The variants in use (after tuning or pinning; 0 for FP64 contexts). */
int pn_get_variants(const pn_ctx* ctx, int* var_x, int* var_y);
/* The 3-D FP32 theta sweep's own shape: var_z 0..3 as var_x (what pn_set_variants sets it
 * to), or 4 = (3, 4, 2), 192-cell windows, which run two theta pencils of up to 84 cells with
 * their halos per wavefront (refused with PN_ERR_ARG for a viscous disk, which keeps 128-cell
 * theta windows).  pn_tune picks it apart from the radial shape.  0 for FP64 and 2-D
 * contexts. */
int pn_set_variant_z(pn_ctx* ctx, int var_z);
int pn_get_variant_z(const pn_ctx* ctx, int* var_z);
/* FP32 radial and theta sweeps: on (1, the default) the last window of a pencil, when it
 * holds at most 32 K cells (K cells per lane), runs two lab columns per wavefront, one in
 * each half; off (0) it runs one, like every other window.  The results are the same either
 * way; the switch exists to show that.  A batch's disks follow their own contexts' setting
 * (pn_batch_disk).  No counterpart in the reference. */
int pn_set_packing(pn_ctx* ctx, int on);

/* The achievable HBM bandwidth: a plain device copy of nbytes (a multiple of 16, at least
 * 1 MiB) launched reps times back to back after one warm-up, timed with HIP events on a
 * stream of its own; *gbs = 2 nbytes reps / time in GB/s (read + write).  Allocates and
 * frees its two buffers.  No counterpart in the reference; the measurements quote kernel
 * bandwidth against it. */
int pn_copy_bandwidth(size_t nbytes, int reps, double* gbs);

/* One application of the damping zones (kill_wave, kill_wave.cu:3-43) with time step dt to
 * the current state, whatever kill_flag says; pn_run applies it after every step's sweeps
 * when kill_flag is 1.  Both precisions. */
int pn_damp_edges(pn_ctx* ctx, double dt);

/* Mass and angular-momentum totals in FP64 with compensated summation. */
int pn_totals(pn_ctx* ctx, double* mass, double* angmom);

/* pn_totals without any allocation: the same kernels, with their per-block partials and the
 * two sums in the library's file-scope device scratch (8192 + 8 doubles per device, held
 * until the process ends) instead of the context's buffers.  No counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Calls no hipMalloc / hipFree.  The
 * reduction runs at most 4096 blocks of two partials, so every grid fits the scratch.  Every
 * call on a device shares that device's scratch: calls that overlap on one device (two host
 * threads, two streams) race. */
int pn_totals_noalloc(pn_ctx* ctx, double* mass, double* angmom);

/* ---- batches of independent disks ---------------------------------------
 * No counterpart in the reference, which advances one disk per process.  Each disk of a
 * batch is an ordinary context (reach it with pn_batch_disk for pn_load_rings,
 * pn_store_rings, pn_totals, snapshots) with its own dt, planet and frame; per step the
 * batch issues one launch for every disk's pre-step, one persistent launch per sweep over
 * the work items of all disks, and one launch for every disk's post-step -- the per-disk
 * operations of main.cu:343-387 exactly as pn_run performs them. */
typedef struct pn_batch pn_batch;

/* n disks of grids[d] / params[d].  precision must be PN_FP32; every disk 2-D, isothermal
 * (EOS 0) or with the energy-consistent adiabatic step (EOS 3), viscous or not, with or without
 * damping zones, with its own radial boundary codes; its planet fixed, growing or moving under
 * the disk's forces (plnt_flag 0..2).  The disks share the closure and the frame (eos 3 or not,
 * bary_flag), which select the kernels of a launch (PN_ERR_UNSUPPORTED otherwise).  A disk with
 * a moving planet or damping zones takes its CFL pass before each step, as pn_run does.  On
 * failure *out still holds a batch whose pn_batch_last_error explains it; pn_batch_destroy it. */
int pn_batch_create(int n, const pn_grid* grids, const pn_params* params, int precision, pn_batch** out);
int pn_batch_destroy(pn_batch* b);
const char* pn_batch_last_error(const pn_batch* b);
int pn_batch_size(const pn_batch* b);
/* The context of disk d (owned by the batch), or NULL. */
pn_ctx* pn_batch_disk(pn_batch* b, int d);
/* pn_run_begin for every disk; states: pn_batch_size records. */
int pn_batch_run_begin(pn_batch* b, pn_run_state* states);
/* pn_run_restart for every disk (each loaded from its picture startat with pn_load_snapshot
 * on pn_batch_disk first); states: pn_batch_size records. */
int pn_batch_run_restart(pn_batch* b, int startat, pn_run_state* states);
/* Advance every disk nsteps steps (states in and out as pn_run's); graph != 0 replays a
 * captured step.  Tunes the sweep shapes on the first call unless pinned.  With land set in
 * a disk's state, that disk's steps follow the endtime / tmovie landing of main.cu:352-366
 * exactly as pn_run applies it, and the call returns after the first step in which any disk
 * landed (its state's event says which; every disk has then taken the same number of steps).
 * A disk that reached endtime must leave the batch: a further call with its land set fails
 * for lack of time left. */
int pn_batch_run(pn_batch* b, int nsteps, int graph, pn_run_state* states);
/* Time every sweep variant over the whole batch (state unchanged) and keep the fastest;
 * pin them instead with pn_batch_set_variants (same indices as pn_set_variants). */
int pn_batch_tune(pn_batch* b, int nrep, int* var_x, int* var_y);
int pn_batch_set_variants(pn_batch* b, int var_x, int var_y);
/* The variants in use (after tuning or pinning). */
int pn_batch_get_variants(const pn_batch* b, int* var_x, int* var_y);
/* Mean ms over nrep launches of {pre-step, r sweep, phi sweep, post-step} of the whole
 * batch, timed with HIP events on the batch's stream.  The state advances. */
int pn_batch_time_kernels(pn_batch* b, int nrep, double* ms4);

/* ---- snapshots and output files (output.cu:48-120, global.cu:27-38, main.cu:113-124, 246-262) ---- */

/* Write the imax*kmax rings of ctx as one raw binary_<label>_<NNNNN> file
 * (no header). */
int pn_save_snapshot(pn_ctx* ctx, const char* path);
/* Read a raw snapshot back into ctx (restart, main.cu:246-262); ghost rings
 * and val must already be loaded.  PN_ERR_IO on a missing or short file. */
int pn_load_snapshot(pn_ctx* ctx, const char* path);
/* The run label (create_label, output.cu:57-88) that names every output file: the grid
 * sizes, the c_s scale, the planet, the radial bounds, viscosity, EOS and frame tags, each
 * number through the reference's int conversion.  EOS 3 is tagged _ad3.  Writes the
 * NUL-terminated label into buf[n]; PN_ERR_ARG if it does not fit. */
int pn_label(const pn_grid* g, const pn_params* p, char* buf, size_t n);
/* frame_num (global.cu:27-38): num as five zero-padded digits ("00042"), the suffix of
 * binary_<label>_ snapshot names (main.cu:124, 313, 403).  PN_ERR_ARG outside 0..99999,
 * where the reference returns an empty suffix. */
int pn_frame_name(int num, char* buf, size_t n);
/* write_para_file (output.cu:90-107): the para_<label>.dat header the reader parses
 * (read/post.cu:65-103), one value per line in default stream formatting; nDev is 1. */
int pn_write_para(const pn_grid* g, const pn_params* p, const char* mainp, const char* label, const char* path);
/* write_initial_file (output.cu:109-120): "xa dx" of each radial cell, one per line. */
int pn_write_initial(const pn_grid* g, const pn_params* p, const char* path);

#ifdef __cplusplus
}
#endif

#endif /* PENGUIN_H */
