/*
 * echo_amd.h - C-ABI of the MI355X (gfx950) per-timestep grid update for
 * ECHO-QGP (reference: gabriele-inghirami/echo-qgp-mhd-devel).
 *
 * The reference has no FFI: its operator API is a set of Fortran module
 * procedures acting on module arrays.  Each entry point below replaces one
 * of them; the comment names the reference procedure (file:line).  Only
 * plain C types cross the boundary (int32_t, double, pointers, sizes), so a
 * Fortran caller binds it with iso_c_binding.
 *
 * Arrays that cross the boundary use the reference's Fortran layout:
 * a(ix, iy, iz, k) column-major, x fastest, variable slowest, FP64, i.e. the
 * C index ((k*nz + iz)*ny + iy)*nx + ix.  State lives on the device between
 * calls; host copies are current only after echo_download_state.
 *
 * Every function returns 0 on success or a positive error code: the
 * reference's own codes where one applies (777/107 superluminal in
 * con2prim, 333 in the flux, 121-138 EoS out of table range) and
 * ECHO_E_* below for errors of this library.  echo_last_error() gives a
 * message for the last failure in the calling thread.
 */
#ifndef ECHO_AMD_H
#define ECHO_AMD_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ECHO_E_ARG     1001  /* invalid argument / unsupported option */
#define ECHO_E_HIP     1002  /* HIP runtime error */
#define ECHO_E_NOMEM   1003  /* device allocation failed */
#define ECHO_E_CRASHED 1004  /* the run crashed earlier (run_crashed) */
#define ECHO_E_FREEZEOUT 1005 /* superluminal velocity on a freeze-out element
                                 (set_fo_*: "you have to fix this", exit(1)) */

/* Run configuration.  Field meaning and units follow the reference
 * param.dat keys parsed in common.f90:874-1215 (key in brackets).  The
 * layout is 16 int32 followed by 19 double, no padding. */
typedef struct echo_config {
  int32_t nx, ny, nz;          /* [NX, NY, NZ] */
  int32_t coordinates;         /* [COORD] 1 Minkowski, 2 Bjorken (Milne) */
  int32_t viscosity;           /* [VISCOUS] 0/1 */
  int32_t bulk;                /* [BULK] 0/1 */
  int32_t mhd;                 /* [MHD] 0 hydro, 1 ideal MHD, 2 resistive
                                  MHD (needs DIVCLEAN; with VISCOUS the
                                  viscous resistive system, nv 19).  The
                                  reference's resistive path does not run
                                  (system.f90:1160-1162); this one is a
                                  design of its own */
  int32_t divclean;            /* [DIVCLEAN] */
  int32_t eqstate;             /* [EQ_STATE] 1 p=e/3, 2 ideal gas, 3 table */
  int32_t recal;               /* [REC_ALGO] 1 TVD2 2 CENO3 3 WENO3 4 WENO5
                                  5 PPM4 6 MPE3 7 MPE5 */
  int32_t nrk;                 /* [INT_ALGO] 2 RK2, 3 RK3, 4 IMEX-SSP3(4,3,3)
                                  (MHD grids, as common.f90:2076-2083) */
  int32_t boundary_cond;       /* [BOUND_CON] 0 periodic, 2 outflow */
  int32_t algo_solver;         /* [ALG_SOLV] MHD con2prim: 4 ideal gas
                                  (EQ_STATE 2), 5 p=e/3 (EQ_STATE 1), 6 p=e/3
                                  Minkowski quartic */
  int32_t obtained_zz;         /* [TRACE_IMP] 0 'no', 1 'zz' */
  int32_t init_type;           /* [INIT_TYPE] (selects viscous tau_pi rules) */
  int32_t derivatives_out;     /* [derivativ] must be 0: out_derivatives is not built */
  double x1, x2, y1, y2, z1, z2; /* [XMIN..ZMAX] */
  double tmin;                 /* [TSTART] */
  double cfl;                  /* [CFL] */
  double maxdt;                /* [MAXDT] */
  double maxspeed;             /* [MAXSPEED] */
  double eta_over_s;           /* [ETA_S] */
  double tau_pi_coeff;         /* [TAU_FACT] */
  double enezero;              /* [ENEZERO] */
  double glm_alpha;            /* [GLM_PARAM] */
  double sigma_el;             /* [SIGMA_EL] */
  double sigma_chi;            /* [SIGMA_CHI] */
  double dtout;                /* [DTOUT] */
  double smooth_temp;          /* [CUT_TEMP] */
  double tmax;                 /* [TSTOP] */
} echo_config;

typedef struct echo_grid echo_grid;   /* one event's grid on one device */

/* Set-up of one grid: replaces echo.f90:73-109 (system_setup,
 * parallel_grid_setup, common_alloc, common_grid, set_eos_pointers,
 * read_data_from_tabulated_file, find_przero, evolve_alloc).
 * eos_rows: for eqstate 3, the table rows as stored in the reference's
 * eos_data files (T, e/T^4, p/T^4, c_s^2), n_rows of them, converted as
 * eos.f90:323-327 does; NULL otherwise.  device: HIP device ordinal. */
int echo_create(const echo_config *cfg, const double *eos_rows, int32_t n_rows,
                int32_t device, echo_grid **out);
void echo_destroy(echo_grid *g);

/* Number of evolved variables nv (system_setup, system.f90:67-128). */
int echo_nv(const echo_grid *g);

/* Initial state from primitives: the per-cell system_prim2cons followed by
 * system_cons2prim of the reference initialisers (init.f90:2937-2958).
 * v: nv*nx*ny*nz doubles (Fortran layout), host memory. */
int echo_init_from_prim(echo_grid *g, const double *v);

/* Glauber Monte-Carlo events, nucleus-nucleus collisions: generate_events
 * (glaubermc.f90:80-635) with the project's random numbers (xoshiro256**
 * seeded by splitmix64; the reference's gfortran rand() stream is not
 * reproduced).  Woods-Saxon nucleons by rejection on r < 12.6 fm, each
 * nucleus moved to its transverse centre of mass, impact parameter with
 * p(b) ~ b on [0, bmax] or fixed, collisions at d^2 <= sigma_NN / pi, events
 * with a collision and at least min_participants participants.  Writes
 * summary.dat, partcoll.bin (participants, (x, y, type) records of 8 + 8 + 4
 * bytes, 0 proton 1 neutron, A then B), spectators.bin (same records) and
 * collisions.bin ((x, y) records) into dir, events one after the other;
 * *n_events: the events written. */
typedef struct echo_glauber_sampler {
  int32_t A;                 /* [NUCLEUS] mass number of both nuclei (<= 300) */
  int32_t Z;                 /* protons: nucleons 1..Z of a nucleus */
  int32_t nconf;             /* [NCONF] nuclear configurations */
  int32_t nbcoll;            /* [NBCOLL] impact parameters tried per configuration */
  int32_t min_participants;  /* [MIN_PART] (reference default 18) */
  int32_t fixed_b;           /* 1: b = bimpact, one per configuration */
  double radius, diffuse;    /* Woods-Saxon R, a (fm), nuclear_data.dat */
  double rho0;               /* Woods-Saxon normalisation roze (cancels out) */
  double sigma_nn;           /* [SIGMA_IN] inelastic NN cross section, mb */
  double bimpact;            /* [B] fm, with fixed_b */
  double bmax;               /* fm (the reference uses 20 for AA) */
} echo_glauber_sampler;
int echo_glauber_generate(const echo_glauber_sampler *s, uint64_t seed, const char *dir,
                          int32_t *n_events);

/* One event of dir (1-based iev), as generate_energy_density_profile and
 * generate_B_field_profile read it (glaubermc.f90:760-790, 1009-1058):
 * participants of A then of B with their types, spectators likewise,
 * collisions.  The caller's arrays hold cap_part, cap_spec, cap_coll (x, y)
 * pairs (and types); ECHO_E_ARG when an event does not fit. */
typedef struct echo_glauber_event {
  double b;
  int32_t n_part_a, n_part_b, n_spec_a, n_spec_b, n_coll;
  double *part_xy;     /* 2 * cap_part */
  int32_t *part_type;  /* cap_part */
  double *spec_xy;     /* 2 * cap_spec */
  int32_t *spec_type;  /* cap_spec */
  double *coll_xy;     /* 2 * cap_coll */
} echo_glauber_event;
int echo_glauber_read_event(const char *dir, int32_t iev, echo_glauber_event *ev, int32_t cap_part,
                            int32_t cap_spec, int32_t cap_coll);

/* Glauber Monte-Carlo initial state computed on the device: the energy
 * deposition of generate_energy_density_profile (glaubermc.f90:686-842,
 * nucleus-nucleus branch) from given participant and collision positions,
 * the primitives of the Glauber-MC initialisation (init.f90:2662-2730:
 * rho = 1, v = 0, p = eos_pressure(rhocenter, e + enezero)), for viscous
 * runs the Navier-Stokes stresses of viscous_initio (viscous.f90:101-136),
 * then prim2cons (+ con2prim) as echo_init_from_prim.  part_a, part_b, coll:
 * (x, y) pairs in fm (host memory), at most 4096 sources in total, as
 * echo_glauber_read_event gives them. */
typedef struct echo_glauber_params {
  double kappa;       /* [KAPPA] */
  double ah;          /* [AH] weight of the binary collisions */
  double sig_mc;      /* [SIG] Gaussian width of a source (fm) */
  double ybeam;       /* beam rapidity log(sqrt(s_NN)/m_N), glaubermc.f90:741 */
  double deta;        /* [DETA] width of the eta plateau */
  double sigeta;      /* [SIGETA] Gaussian width of the eta tails */
  double rhocenter;   /* [RHOCENTER] */
} echo_glauber_params;
int echo_init_glauber(echo_grid *g, const echo_glauber_params *p, const double *part_a, int32_t n_a,
                      const double *part_b, int32_t n_b, const double *coll, int32_t n_coll);

/* The same for an MHD grid (ideal or resistive), with the initial magnetic field of
 * generate_B_field_profile (glaubermc.f90:877-1174, called from init.f90:
 * 2796-2811 when no B-field file is given; one event, avg_events = 0): the field of the protons among
 * the spectators and participants of nuclei A and B (participants weighted
 * by pw) in a medium of the grid's sigma_el and sigma_chi, added to B^x,
 * B^y, B^eta before prim2cons (on resistive grids E and psi start at zero,
 * as init.f90:2732-2832 fills only kbx..kbz).  lists[4]: spectators A, participants A,
 * spectators B, participants B; each is read up to its first neutron, as
 * the reference does (its files list protons first); at most 2048 protons. */
typedef struct echo_bfield_params {
  double gamma_col;       /* beam Lorentz factor, rads/(2 am) (common.f90:2261) */
  double b_amplify;       /* [B_amp_fac] B_amplify_factor */
  double pw;              /* [PW.......] weight of the participants */
  int32_t magfield_type;  /* [MAGFIELDT] 1 classical + chiral, 2 classical, 3 chiral */
} echo_bfield_params;
typedef struct echo_nucleons {
  const double *xy;       /* (x, y) pairs in fm, host memory */
  const int32_t *type;    /* 0 proton, 1 neutron */
  int32_t n;
} echo_nucleons;
int echo_init_glauber_mhd(echo_grid *g, const echo_glauber_params *p, const double *part_a, int32_t n_a,
                          const double *part_b, int32_t n_b, const double *coll, int32_t n_coll,
                          const echo_bfield_params *bp, const echo_nucleons *lists);

/* Upload / download primitive and conserved state (host memory, Fortran
 * layout).  Replaces the reference's direct access to module arrays v, u
 * (common.f90:203-214). */
int echo_upload_state(echo_grid *g, const double *v, const double *u);
int echo_download_state(echo_grid *g, double *v, double *u);

/* nsteps full time steps: the body of the reference main loop echo.f90:
 * 185-241 (vnewstep/vold bookkeeping, evolve_main for irk = 1..nrk, time
 * and metric update), all on the device.  Returns 0 or the error code of a
 * crashed step (the reference's errcode). */
int echo_steps(echo_grid *g, int32_t nsteps);

/* One RK stage: evolve_main(irk, dt), evolve.f90:114-298.  Stages must be
 * called in order 1..nrk; echo_end_step then does the bookkeeping of
 * echo.f90:192-241.  echo_steps is the fused equivalent. */
int echo_rk_stage(echo_grid *g, int32_t irk);
int echo_end_step(echo_grid *g);

/* Scalars of the run (reads device state):
 * s[0] t  s[1] dt  s[2] timeold  s[3] timeinterval  s[4] g_cov_3_old
 * s[5] glm_ch_old  s[6] tout  s[7] iterations  s[8] run_crashed  s[9] errcode */
int echo_get_scalars(echo_grid *g, double *s);

/* Number of cells repaired since set-up by the con2prim repair of
 * evolve_u2v (evolve.f90:795-842): cells whose cons2prim failed and that
 * were rebuilt from the average of their face neighbours. */
int echo_get_repairs(echo_grid *g, int64_t *n);

/* Bulk viscosity with zeta/s peaked at the transition temperature:
 * zeta = hbar zs_max exp(-(T - t_c)^2 / (2 width^2)) s, T and t_c, width in
 * GeV, in place of the reference's zeta = 2 eta (1/3 - c_s^2)
 * (get_viscous_parameters, viscous.f90:186-190), in the sources and in
 * echo_init_glauber's Navier-Stokes Pi.  No counterpart in the reference.
 * Needs VISCOUS and BULK; call between steps. */
int echo_set_bulk_peak(echo_grid *g, double zs_max, double t_c, double width);

/* Next output time tout (echo.f90:172, 243-247); the dt clamps of
 * evolve.f90:179-188 use it. */
int echo_set_tout(echo_grid *g, double tout);

/* Output frame of out_vars (out.f90:232-591): the reference's stream file
 * format -- t, then the fields selected in sel[0..27] (the flags of
 * type out_type, common.f90:122-151: density, vx, vy, vz, pressure,
 * energy_density, temperature, entropy_density, bulk, pitt, pitx, pity,
 * pitz, pixy, pixz, piyz, pixx, piyy, pizz, v0, bx, by, bz, ex, ey, ez, glm,
 * rc) as nx*ny*nz arrays, FP64 (precision 8) or FP32 (4) -- written to
 * <dir>/out<iout, 4 digits>.dat, with "iout t" appended to <dir>/time.dat.
 * The state and the derived fields are copied on the device when the call
 * is made; the device-to-host copy and the file write then run in the
 * background while the grid keeps stepping.  echo_wait_frames waits for the
 * frame in flight (a new echo_write_frame waits for the previous one).
 * On resistive grids ex, ey, ez are the evolved E as held (out.f90:404-413),
 * on ideal-MHD grids -v x B (414-429).
 * rc is not supported; pizz needs TRACE_IMP 'zz' and a pi^{t mu} field. */
int echo_write_frame(echo_grid *g, const int32_t *sel, int32_t precision, int32_t iout,
                     const char *dir);
int echo_wait_frames(echo_grid *g);

/* Stop criterion check_temp (out.f90:1487-1557; echo.f90:189-190): with
 * temp_end > 0, every step starts by taking the maximum over the cells of
 * the temperature (freeze_type 0) or the energy density (freeze_type 1, where
 * temp_end is an energy density); below temp_end the run stops: that step and
 * all later ones do nothing (t, iter and the state stay), as the reference
 * leaves its time loop.  temp_end <= 0 disables the check (the default).
 * echo_get_stop: stopped flag and the maximum of the last check. */
int echo_set_stop(echo_grid *g, double temp_end, int32_t freeze_type);
int echo_get_stop(echo_grid *g, int32_t *stopped, double *max_value);

/* Restart dump: out_dump / out_restart (out.f90:1313-1483) in the
 * reference's dump.dat format (a Fortran stream file: 4-byte iout, t,
 * timeold, timeinterval, v, u, vold, uold -- components krh..kvold_end,
 * up to p for hydro (5), up to B^eta for MHD (8; 15 on the viscous
 * resistive grid, whose B follows the viscous variables) --, g_cov_3_old,
 * and with echo_fo_setup active:
 * thyp, time_arr, a primitive slice, field(2,:,:,:)).
 * echo_set_dump(g, 1) keeps vold / uold on the device (one copy of
 * 2 * kvold_end fields per step, at the step's start, as echo.f90:192-201);
 * call it before the steps to be dumped (until a step has run, vold / uold
 * are the current state, as init.f90:3271-3272).  echo_dump writes the
 * file between steps; echo_restart reads one into the grid (t, timeold,
 * timeinterval, g_cov_3_old, v, u, vold; tout = t + dtout as
 * echo.f90:164-172) and returns iout and thyp.  In the freeze-out part the
 * primitive slice is the one the next echo_fo_slice compares against (the
 * reference stores its older slice there, which its own restart cannot
 * use: the next out_hyp overwrites it with primitives, which the file does
 * not hold). */
int echo_set_dump(echo_grid *g, int32_t on);
int echo_dump(echo_grid *g, const char *path, int32_t iout, double thyp);
int echo_restart(echo_grid *g, const char *path, int32_t *iout, double *thyp);

/* Freeze-out hypersurface: the out_freeze path of the reference.
 * echo_fo_setup replaces allocate_hypersurface_arrays
 * (hypersurface.f90:93-168) and get_hysufile_ready (1137-1200): freeze-out
 * on the temperature (freeze_type 0) or the energy density (1) at
 * freeze_value; hdr (may be null) receives the header's temperature and
 * energy density.  echo_fo_slice replaces out_hyp (out.f90:1568-1645) with
 * find_hypersurface_{ideal,visco,mhd} (hypersurface.f90:172-229, 466-521,
 * 775-830): the first call stores the current state as a slice; every
 * later call finds the crossings of freeze_value between the stored slice
 * and the current state (along tau in each cell, along x, y, eta between
 * neighbours of the older slice), in the reference's order, and stores the
 * current state in turn.  Each element is one row of hypersurface.txt:
 * echo_fo_width values -- 13 for ideal hydro (tau, x, y, eta, rho, v^x,
 * v^y, v^eta, p, dSigma_tau, dSigma_x, dSigma_y, dSigma_eta), 25 for
 * viscous hydro (+ Pi, pi^xy, pi^xeta, pi^yeta, pi^xx, pi^yy, pi^etaeta,
 * pi^tautau, pi^taux, pi^tauy, pi^taueta, 1/(T^2 (e+p))), 16 for MHD
 * (+ B^x, B^y, B^eta).  echo_fo_elements copies the n elements of the last
 * slice (row-major, n * width doubles).  echo_fo_write writes
 * hypersurface.txt as print_hypersurface_* does: header = 1 creates the
 * file with its header line, header = 0 appends the last slice (nothing
 * after the first, storing, slice).  Errors:
 * an EoS code (nothing changes), ECHO_E_FREEZEOUT (v^2 > 1 on an element,
 * where the reference exits). */
int echo_fo_setup(echo_grid *g, int32_t freeze_type, double freeze_value, double *hdr);
int echo_fo_slice(echo_grid *g, int64_t *n_elements);
int echo_fo_width(const echo_grid *g);
int echo_fo_elements(echo_grid *g, double *out, int64_t n);
int echo_fo_write(echo_grid *g, const char *path, int32_t header);

/* Batched con2prim: system_cons2prim (system.f90:423) on n independent
 * states at the grid's current metric.  u: nv*n conserved (variable-major,
 * u[k*n + i]); v: nv*n primitives in/out (in: guess, out: solution);
 * status: n error codes (0 = converged).  Host pointers. */
int echo_cons2prim_batch(echo_grid *g, const double *u, double *v, int32_t *status,
                         int64_t n);

/* Tile shape of the flux sweeps and its autotuner (no counterpart in the
 * reference, whose sweeps are pencil loops, evolve.f90:302-685).  A y or z
 * sweep with `rows` = 2 or 4 runs blocks of 256/rows faces along x by `rows`
 * faces along the sweep (WENO5 and PPM4); results are identical for every
 * shape.  echo_tune times each shape on the current state (reps launches
 * each) and keeps the fastest; ms (optional, 9 doubles): per-launch times,
 * [3*sweep + {0,1,2}] for rows {1,2,4}. */
int echo_set_tile(echo_grid *g, int32_t sweep, int32_t rows);
int echo_get_tile(const echo_grid *g, int32_t *rows);
int echo_tune(echo_grid *g, int32_t reps, double *ms);

/* ---- event-by-event ensemble (no counterpart in the reference, which runs
 * one event per process, echo.f90:134-280) ---------------------------------
 * n_events independent grids (configurations may differ in size, range,
 * physics) on one device, sharing one stream and one pool of face buffers
 * sized for the largest event.  echo_ensemble_grid gives event i's grid for
 * the per-grid calls (init, upload, download, scalars; not echo_destroy).
 * echo_ensemble_steps advances every event by nsteps time steps, each step
 * one replay of a captured graph holding all events' kernels; *n_crashed
 * counts crashed events (their scalars give the error codes). */
typedef struct echo_ensemble echo_ensemble;
int echo_ensemble_create(const echo_config *cfgs, int32_t n_events, const double *eos_rows,
                         int32_t n_rows, int32_t device, echo_ensemble **out);
void echo_ensemble_destroy(echo_ensemble *e);
int echo_ensemble_size(const echo_ensemble *e);
echo_grid *echo_ensemble_grid(echo_ensemble *e, int32_t i);
int echo_ensemble_tune(echo_ensemble *e, int32_t reps);
int echo_ensemble_steps(echo_ensemble *e, int32_t nsteps, int32_t *n_crashed);
/* wall time (HIP events) of the last echo_ensemble_steps call, ms */
int echo_ensemble_last_ms(const echo_ensemble *e, double *ms);

/* Batched con2prim on device-resident states (device pointers, same layout
 * as echo_cons2prim_batch): the plain solver of the grid's system. */
int echo_cons2prim_batch_dev(echo_grid *g, const double *d_u, double *d_v, int32_t *d_status,
                             int64_t n);

/* ---- learned con2prim (no counterpart in the reference) ----------------
 * An 8 -> 64 -> 64 -> 64 -> 5 network (half-precision MFMA, FP32 accumulation)
 * supplies the initial guess (v^2, W) of the ideal-MHD p = e/3 inversion
 * that the reference does with algorithm 5 of system_cons2prim
 * (system.f90:945-998, rtsafe + solver8); `iters` FP64 Newton iterations on
 * the two equations of solver8 refine it; a state is accepted when the
 * Newton correction at the final point is below tol relative to 1 - v^2 and
 * to W, else the plain solver (from the guess in d_v) handles it.
 * Parameters: ECHO_MLP_NPARAM floats, W1[64][8] b1[64] W2[64][64] b2[64]
 * W3[64][64] b3[64] W4[5][64] b4[5] (row-major [out][in]). */
#define ECHO_MLP_NPARAM 9221
typedef struct echo_mlp echo_mlp;

typedef struct echo_mlp_train_cfg {
  int32_t steps;         /* optimiser steps */
  int32_t batch;         /* states per step (consecutive slices of the data) */
  float lr, lr_final;    /* Adam step size, cosine schedule lr -> lr_final */
  float beta1, beta2, eps;
} echo_mlp_train_cfg;

int echo_mlp_create(int32_t device, echo_mlp **out);
void echo_mlp_destroy(echo_mlp *m);
int echo_mlp_set_params(echo_mlp *m, const float *p, int64_t n);
int echo_mlp_get_params(echo_mlp *m, float *p, int64_t n);
/* Network outputs y[5][n] (device) for n conserved states d_u[8][n]
 * (Minkowski metric). */
int echo_mlp_forward(echo_mlp *m, echo_grid *g, const double *d_u, int64_t n, float *d_y);
/* Training loss and its gradient (host outputs) over n labelled states
 * (d_u conserved, d_v primitives of the plain solver, [8][n], Minkowski). */
int echo_mlp_grad(echo_mlp *m, echo_grid *g, const double *d_u, const double *d_v, int64_t n,
                  float *grad, double *loss);
/* Adam training on the device; loss_log (host, optional): cfg->steps
 * per-step losses. */
int echo_mlp_train(echo_mlp *m, echo_grid *g, const double *d_u, const double *d_v, int64_t n,
                   const echo_mlp_train_cfg *cfg, float *loss_log);
/* The learned inversion on n states at the grid's current metric (device
 * pointers; d_v: guess in, primitives out).  The grid must be configured
 * for ideal MHD with p = e/3 (mhd 1, algo_solver 5, eqstate 1).
 * *n_fallback: states handed to the plain solver. */
int echo_c2p_surrogate(echo_grid *g, echo_mlp *m, const double *d_u, double *d_v,
                       int32_t *d_status, int64_t n, int32_t iters, double tol,
                       int64_t *n_fallback);

/* echo_c2p_surrogate without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The fallback
 * counter and list live in a static device array of the library, one per
 * device (4 + 4 * ECHO_NOALLOC_MAX_STATES bytes, about 16 MB, held until the
 * process ends); n above ECHO_NOALLOC_MAX_STATES is rejected with ECHO_E_ARG.
 * Results equal echo_c2p_surrogate's.  Every call on a device shares that
 * device's array: calls that overlap on one device race. */
#define ECHO_NOALLOC_MAX_STATES (1 << 22)
int echo_c2p_surrogate_noalloc(echo_grid *g, echo_mlp *m, const double *d_u, double *d_v,
                               int32_t *d_status, int64_t n, int32_t iters, double tol,
                               int64_t *n_fallback);

/* The learned inversion inside the grid's own steps: replaces the con2prim
 * of each RK stage (evolve_u2v, evolve.f90:769-868, algorithm 5 of
 * system_cons2prim, system.f90:945-998) with the network's guess, iters
 * FP64 Newton iterations and the gate of tolerance tol; the states the gate
 * rejects go to algorithm 5 from the stage-start primitives, as without the
 * network, and cells that fail there are repaired as usual.  Ideal MHD with
 * p = e/3 only (mhd 1, algo_solver 5, eqstate 1).  m = NULL detaches the
 * network.  The grid keeps the pointer: detach it (or destroy the grid)
 * before echo_mlp_destroy.  Allocates a fallback list of N + 1 int32 on the
 * first attach.  Not between echo_rk_stage calls of one step. */
int echo_set_c2p_net(echo_grid *g, echo_mlp *m, int32_t iters, double tol);
/* States of the grid's steps that the gate left to the plain solver, since
 * set-up. */
int echo_get_c2p_fallbacks(echo_grid *g, int64_t *n);

/* Kernel timing in milliseconds, measured with HIP events on the grid's
 * stream when enabled with echo_set_timing(g, 1).  After echo_steps (per
 * launch): t[0] flux x, t[1] flux y, t[2] flux z, t[3] cell update, means
 * over the stages that sweep (all RK stages; IMEX stages 2-4, the first
 * IMEX stage's implicit solves are in no mean).  After
 * echo_cons2prim_batch_dev: t[0] the solver kernel.  After
 * echo_c2p_surrogate: t[0] network + refinement kernel, t[1] fallback. */
int echo_set_timing(echo_grid *g, int32_t on);
int echo_get_timing(echo_grid *g, double *t);

/* Device memory for callers without a HIP runtime of their own (the
 * _dev and network entry points take such pointers).  Copies are
 * synchronous. */
int echo_dev_alloc(int32_t device, int64_t bytes, void **ptr);
int echo_dev_free(int32_t device, void *ptr);
int echo_dev_upload(int32_t device, void *dst, const void *src, int64_t bytes);
int echo_dev_download(int32_t device, void *dst, const void *src, int64_t bytes);

const char *echo_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
