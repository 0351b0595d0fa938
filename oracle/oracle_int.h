/* ORACLE internals - test infrastructure, not product code (see oracle.h). */
#ifndef ECHO_ORACLE_INT_H
#define ECHO_ORACLE_INT_H
#include <math.h>
#include <stddef.h>

#define NGC 3          /* ghost cells, common.f90:184 */
#define NDERIV 20      /* common.f90:79 */

/* primitive / conserved slots, common.f90:63 (0-based here) */
enum { KRH = 0, KVX = 1, KVY = 2, KVZ = 3, KPR = 4, KPIBU = 5, KPIXY = 6,
       KPIXZ = 7, KPIYZ = 8, KPIXX = 9, KPIYY = 10, KPIZZ = 11 };
/* deriv slots, common.f90:70-72 (0-based here) */
enum { DTX = 0, DTY, DTZ, DXX, DXY, DXZ, DYX, DYY, DYZ, DZX, DZY, DZZ,
       DTT, DXT, DYT, DZT, DTET, DTEX, DTEY, DTEZ };
/* shear components, viscous.f90:43 (0-based here) */
enum { SXX = 0, SYY = 1, SZZ = 2, SXY = 3, SXZ = 4, SYZ = 5 };

enum { MINKOWSKI = 1, BJORKEN = 2 };
enum { RK2 = 2, RK3 = 3, SSP = 4 };
enum { GUBSER_VISCOUS = 5, SHEAR_VISCOUS_1D = 2 };

#define MAXNV 20

typedef struct {
  /* configuration */
  int nx, ny, nz, coordinates, viscosity, bulkvis, mhd, rmhd, divclean;
  int eqstate, recal, nrk, boundary_cond, algo_solver, obtained_zz;
  int init_type, derivatives_out;
  double x1, x2, y1, y2, z1, z2, tmin, cfl, maxdt, maxspeed, eta_over_s;
  double tau_pi_coeff, enezero, glm_alpha, sigma_el, sigma_chi, dtout;
  /* bulk viscosity zeta/s peaked at T_c (orc_set_bulk_peak); 0: the
   * reference's zeta = 2 eta (1/3 - c_s^2) */
  int zeta_peak;
  double zs_max, zs_tc, zs_w;
  double smooth_temp, tmax;
  int enable_smooth_viscosity;
  /* system_setup, system.f90:67-128 */
  int nv, nu, nov, nou, kbx, kby, kbz, kex, key, kez, kglm;
  /* grid, common.f90:441-481 */
  double dx, dy, dz, ddx, ddy, ddz;
  double *xc, *yc, *zc;
  /* metric and step bookkeeping */
  double t, timeold, timeinterval, g_cov[3], gp, gm, g_cov0, g_cov_3_old;
  double dt, tout, glm_ch, glm_ch_old, aflux[3];
  int iter, run_crashed, errcode;
  /* EoS */
  double przero;
  int entries;
  double *tab_t, *tab_e, *tab_p, *tab_c;
  double smooth_pressure; int smooth_first;
  /* check_temp stop criterion */
  double temp_end, check_max; int freeze_type, temp_check;
  /* fields */
  double *v, *u, *u0, *du, *du_stiff, *deriv, *vnewstep, *vold, *unewstep, *uold;
  double *du_all, *dus_all;   /* 4 slots each: du, du_stiff of the IMEX stages */
  double imex_adt;            /* alpha dt of the implicit solve in evolve_u2v, 0: none */
  int kvold_end;   /* components krh..kvold_end of vold/uold: 5 (hydro), 8 (MHD) */
  double *ax, *ay, *az;
  /* how often each repair branch fired since orc_setup (REP_* below) */
  long nrep[8];
  /* cells (cell_index) of the first REPLOG u2v repairs since orc_setup */
#define REPLOG 256
  long replog[REPLOG];
  long crash_cell;   /* the cell whose con2prim repair failed, -1: none */
} orc_state;

/* repair branches of evolve_df{x,y,z} (evolve.f90:334-369) and
 * evolve_u2v (evolve.f90:795-842) */
enum { REP_REC_PR = 0,    /* reconstructed p <= 0 replaced by the cell's */
       REP_REC_V = 1,     /* reconstructed v^2 >= 1 replaced by the cell's */
       REP_REC_VMAX = 2,  /* the cell's own v^2 >= 1: scaled to maxspeed */
       REP_U2V = 3,       /* cons2prim failed: neighbour-average repair */
       REP_U2V_V = 4,     /* repaired v^2 > 1 scaled to maxspeed */
       REP_U2V_PR = 5,    /* repaired p < 0 replaced by the neighbour average */
       REP_U2V_FAIL = 6,  /* repair failed: run crashed */
       REP_C2P_UNDEF = 7 };/* an ideal or viscous cons2prim Newton step used
                              eos_pressure outputs before any call of that
                              cons2prim had set them: the reference then
                              reads uninitialised locals (see sys_cons2prim) */

extern orc_state G;

/* EoS (eos.f90) */
int eos_pressure(double rh, double en, double *pr, double *dprdrh, double *dprden);
int eos_energy(double rh, double pr, double *en);
int eos_sound(double rh, double en, double pr, double *cs2);
int get_derived_data(double rh, double pr, double *en, double *temp, double *s);
int eos_load_table(const char *path);
int find_pressure_from_temperature(double temp_input, double *pres);
void fo_release(void);   /* oracle_fo.c: drops the freeze-out slices */
#include <stdio.h>
int fo_dump_part(FILE *f, double thyp);
int fo_restart_part(FILE *f, double *thyp);

/* reconstruction (holib.f90) - f points to element 0 of f(-ngc..n+ngc) */
void rec_pencil(int code, const double *f, double *gl, double *gr, int n);
int rec_ngc(int code);

/* system (system.f90 / viscous.f90) */
int sys_prim2cons(double *v, double *u);
int sys_prim2cons_0(double *v, double *u);
int sys_cons2prim(double *u, double *v);
int sys_flux(double *v, double *u, double *f, double *vf, int dir);
int sys_source(int ix, int iy, int iz, double *v, double *src, double *src_stiff);
/* resistive MHD with or without viscosity (oracle_rmhd.c) */
int rmhd_nfluid(void);
int rmhd_prim2cons(double *v, double *u, int initial);
int rmhd_cons2prim(double *u, double *v);
int rmhd_cons2prim_implicit(double *u, double *v, double adt);
int rmhd_flux(double *v, double *u, double *f, double *vf, int dir);
int rmhd_source(int ix, int iy, int iz, double *v, double *src, double *src_stiff);
void rmhd_stiff(const double *v, double *s3);
double bulk_zeta(double eta, double cs2, double temp, double s);
void get_derived_pi(const double *p, double *pitt, double *pitx, double *pity, double *pitz);
void get_derived_pi_zz(const double *p, double *pitt, double *pitx, double *pity, double *pitz, double *pizz);
void system_metric(void);

static inline size_t cell_index(int ix, int iy, int iz) {
  return ((size_t)iz * G.ny + iy) * G.nx + ix;
}
static inline size_t ncell(void) { return (size_t)G.nx * G.ny * G.nz; }
#define V(ix, iy, iz, k) G.v[(size_t)(k) * ncell() + cell_index(ix, iy, iz)]
#define U(ix, iy, iz, k) G.u[(size_t)(k) * ncell() + cell_index(ix, iy, iz)]
#define DERIV(ix, iy, iz, k) G.deriv[(size_t)(k) * ncell() + cell_index(ix, iy, iz)]

static inline double fmin2(double a, double b) { return a < b ? a : b; }
static inline double fmax2(double a, double b) { return a > b ? a : b; }

#endif
