/*
 * ORACLE - test infrastructure, not product code.
 *
 * A sequential FP64 CPU restatement, in plain C, of ECHO-QGP's per-timestep
 * grid update (reference: gabriele-inghirami/echo-qgp-mhd-devel).  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it (oracle/liboracle.so), and only as the checker or the timed CPU
 * baseline; the HIP product never links it.
 *
 * Every function cites the reference file:line it restates.  The oracle is
 * pinned against the reference Fortran itself (oracle/_ref, built from the
 * reference sources by oracle/build_ref.sh) and against closed forms; the
 * pins live in tests/test_oracle_*.py.
 *
 * Layout: arrays are Fortran order (x fastest, variable slowest), 0-based:
 * a[((k*nz + iz)*ny + iy)*nx + ix].
 */
#ifndef ECHO_ORACLE_H
#define ECHO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Configuration, same meaning as the reference harness arrays
 * (oracle/ref_harness.f90):
 *  ic[0..2] nx,ny,nz   ic[3] coordinates 1/2   ic[4] viscosity
 *  ic[5] bulk          ic[6] mhd 0/1/2         ic[7] divclean
 *  ic[8] eqstate 1/2/3 ic[9] reconstruction    ic[10] nrk 2/3/4
 *  ic[11] boundary_cond 0/2  ic[12] algo_solver  ic[13] obtained zz
 *  ic[14] init_type    ic[15] derivatives_out
 *  rc[0..5] x1,x2,y1,y2,z1,z2  rc[6] tmin  rc[7] cfl  rc[8] maxdt
 *  rc[9] maxspeed  rc[10] eta_over_s  rc[11] tau_pi_coeff  rc[12] enezero
 *  rc[13] glm_alpha rc[14] sigma_el  rc[15] sigma_chi  rc[16] dtout
 *  rc[17] smooth_temp  rc[18] tmax
 */
int  orc_setup(const int *ic, const double *rc, const char *table_path);
void orc_free(void);
int  orc_nv(void);
int  orc_init_from_prim(const double *vin);
void orc_set_state(const double *v, const double *u);
void orc_get_state(double *v, double *u);
void orc_get_deriv(double *d);
void orc_get_scalars(double *scal);
int  orc_steps(int nsteps);
/* bulk viscosity zeta/s = zs_max exp(-(T - t_c)^2 / (2 width^2)) (T in
 * GeV) instead of the reference's 2 eta (1/3 - c_s^2) / s, until the next
 * orc_setup; needs viscosity with bulk */
void orc_set_bulk_peak(double zs_max, double t_c, double width);
/* MAXDT from now on (a test that follows a larger grid's time step) */
void orc_set_maxdt(double maxdt);
/* one RK stage irk (evolve_main) without the step bookkeeping */
int  orc_stage(int irk);
/* evolve_u2v alone on the current v, u */
int  orc_u2v(void);
/* out[0..7]: how often each repair branch fired since orc_setup
 * (REP_* of oracle_int.h) */
void orc_get_repair_counts(long *out);
/* the cells (x fastest) of the first min(n, 256) u2v repairs; returns how
 * many were written */
int  orc_get_repair_cells(long *out, int n);
/* the cell (x fastest) whose con2prim repair failed and crashed the run,
 * -1 if none */
long orc_crash_cell(void);

/* pointwise operators at the current metric */
void orc_set_time(double t);
void orc_rec(int code, int n, int ngc_in, const double *f, double *gl, double *gr);
int  orc_prim2cons(double *v, double *u);
int  orc_cons2prim(double *u, double *v);
int  orc_flux(double *v, int dir, double *u, double *f, double *vf);
int  orc_eos_pressure(double rh, double en, double *out3);
int  orc_eos_energy(double rh, double pr, double *en);
int  orc_eos_sound(double rh, double en, double pr, double *cs2);
int  orc_eos_derived(double rh, double pr, double *out3);
int  orc_eos_table(int n, double *t, double *e, double *p, double *c);
double orc_przero(void);
void orc_set_stop(double temp_end, int freeze_type);
int orc_out_frame(const int *sel, int prec, int iout, const char *path, const char *tpath);
void orc_get_stop(int *stopped, double *max_value);

/* Glauber-MC initial primitives (oracle_init.c) */
int orc_glauber_prim(const double *par, const double *pa, int na, const double *pb, int nb,
                     const double *pc, int nc, double *vout);
int orc_glauber_bfield(const double *par, const double *sa, const int *ta, int nsa, const double *pa,
                       const int *tpa, int npa, const double *sb, const int *tb, int nsb,
                       const double *pb, const int *tpb, int npb, double *v);

/* freeze-out hypersurface (oracle_fo.c) */
int  orc_fo_setup(int freeze_type, double freeze_value, double *hdr);
int  orc_fo_slice(long *n_elements);
int  orc_fo_width(void);
void orc_fo_get(double *out);
int  orc_fo_write(const char *path, int header);

/* restart dump (oracle_out.c) */
int  orc_dump(const char *path, int iout, double thyp);
int  orc_restart(const char *path, int *iout, double *thyp);

#ifdef __cplusplus
}
#endif
#endif
