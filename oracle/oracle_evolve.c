/* ORACLE - test infrastructure, not product code (see oracle.h).
 * Grid, boundaries, directional sweeps, sources, RK stage and time step:
 * restates evolve.f90, work.f90 (HLL), common.f90:441-481 (grid),
 * viscous.f90:56-97 (dts) and echo.f90:185-241 (step bookkeeping).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "oracle_int.h"

orc_state G;

static void free_fields(void) {
  free(G.v); free(G.u); free(G.u0); free(G.du_all); free(G.dus_all);
  free(G.deriv); free(G.vnewstep); free(G.vold); free(G.unewstep); free(G.uold);
  free(G.ax); free(G.ay); free(G.az);
  free(G.xc); free(G.yc); free(G.zc);
  G.v = G.u = G.u0 = G.du = G.du_stiff = G.deriv = G.vnewstep = G.vold = G.unewstep = G.uold = NULL;
  G.du_all = G.dus_all = NULL;
  G.ax = G.ay = G.az = G.xc = G.yc = G.zc = NULL;
  fo_release();
}

void orc_free(void) {
  free_fields();
  free(G.tab_t); free(G.tab_e); free(G.tab_p); free(G.tab_c);
  G.tab_t = G.tab_e = G.tab_p = G.tab_c = NULL;
}

int orc_nv(void) { return G.nv; }

/* system_setup, system.f90:67-128 (HD / viscous / ideal MHD with GLM) */
static void system_setup(void) {
  G.nov = 5; G.nou = 5;
  G.nv = G.viscosity ? 12 : 5;
  G.kbx = G.kby = G.kbz = G.kglm = -1;
  G.kex = G.key = G.kez = -1;
  if (G.mhd && G.rmhd) {
    G.nov += 6; G.nou += 6; G.nv += 6;
    G.kbx = G.nv - 6; G.kby = G.nv - 5; G.kbz = G.nv - 4;
    G.kex = G.nv - 3; G.key = G.nv - 2; G.kez = G.nv - 1;
    if (G.divclean) { G.nov++; G.nou++; G.nv++; G.kglm = G.nv - 1; }
  } else if (G.mhd) {
    G.nov += 3; G.nou += 3; G.nv += 3;
    G.kbx = G.nv - 3; G.kby = G.nv - 2; G.kbz = G.nv - 1;
    if (G.divclean) { G.nov++; G.nou++; G.nv++; G.kglm = G.nv - 1; }
  }
  G.nu = G.nv;
}

/* common_grid, common.f90:441-481 */
static void common_grid(void) {
  G.dx = (G.x2 - G.x1) / G.nx;
  G.dy = (G.y2 - G.y1) / G.ny;
  G.dz = (G.z2 - G.z1) / G.nz;
  G.ddx = G.nx > 1 ? 1. / G.dx : 0.;
  G.ddy = G.ny > 1 ? 1. / G.dy : 0.;
  G.ddz = G.nz > 1 ? 1. / G.dz : 0.;
  G.xc = malloc(sizeof(double) * G.nx);
  G.yc = malloc(sizeof(double) * G.ny);
  G.zc = malloc(sizeof(double) * G.nz);
  for (int i = 1; i <= G.nx; i++) G.xc[i - 1] = G.x1 + (i - 0.5) * G.dx;
  for (int i = 1; i <= G.ny; i++) G.yc[i - 1] = G.y1 + (i - 0.5) * G.dy;
  for (int i = 1; i <= G.nz; i++) G.zc[i - 1] = G.z1 + (i - 0.5) * G.dz;
}

int orc_setup(const int *ic, const double *rc, const char *table_path) {
  orc_free();
  G.nx = ic[0]; G.ny = ic[1]; G.nz = ic[2];
  G.coordinates = ic[3];
  G.viscosity = ic[4] != 0;
  G.bulkvis = (ic[5] != 0) && G.viscosity;
  G.mhd = ic[6] != 0; G.rmhd = ic[6] == 2;
  G.divclean = (ic[7] != 0) && G.mhd;
  G.eqstate = ic[8]; G.recal = ic[9]; G.nrk = ic[10];
  G.boundary_cond = ic[11]; G.algo_solver = ic[12]; G.obtained_zz = ic[13] != 0;
  G.init_type = ic[14]; G.derivatives_out = ic[15] != 0;
  G.x1 = rc[0]; G.x2 = rc[1]; G.y1 = rc[2]; G.y2 = rc[3]; G.z1 = rc[4]; G.z2 = rc[5];
  G.tmin = rc[6]; G.cfl = rc[7]; G.maxdt = rc[8]; G.maxspeed = rc[9];
  G.eta_over_s = rc[10]; G.tau_pi_coeff = rc[11]; G.enezero = rc[12];
  G.zeta_peak = 0; G.zs_max = G.zs_tc = G.zs_w = 0.;
  G.glm_alpha = rc[13]; G.sigma_el = rc[14]; G.sigma_chi = rc[15];
  G.dtout = rc[16]; G.smooth_temp = rc[17]; G.tmax = rc[18];
  G.enable_smooth_viscosity = G.smooth_temp > 0.;
  G.glm_ch_old = 0.5;
  G.aflux[0] = G.aflux[1] = G.aflux[2] = 0.;
  if (G.rmhd) G.glm_ch_old = 1.;   /* resistive systems clean at the speed of light */
  /* the resistive system (oracle_rmhd.c) with or without viscosity, with
   * divergence cleaning; ideal MHD with viscosity does not exist (the
   * reference refuses viscosity with MHD, common.f90:1904-1908) */
  if (G.rmhd && !G.divclean) return -11;
  if (G.mhd && G.viscosity && !G.rmhd) return -11;
  if (G.nrk == SSP && !G.mhd) return -11;   /* common.f90:2076-2083 */
  system_setup();
  common_grid();
  if (G.eqstate == 3) {
    if (!table_path || eos_load_table(table_path)) return -1;
  }
  /* find_przero, eos.f90:192-218 */
  {
    double d1, d2;
    int e = eos_pressure(0., G.enezero, &G.przero, &d1, &d2);
    if (e || G.przero <= 0.) return -2;
  }
  size_t n = ncell();
  G.v = calloc(n * G.nv, sizeof(double));
  G.u = calloc(n * G.nv, sizeof(double));
  G.u0 = calloc(n * G.nv, sizeof(double));
  /* du, du_stiff: one slot per IMEX stage (evolve.f90:49-79), one for RK */
  G.du_all = calloc(n * G.nv * 4, sizeof(double));
  G.dus_all = calloc(n * G.nv * 4, sizeof(double));
  G.du = G.du_all;
  G.du_stiff = G.dus_all;
  G.deriv = calloc(n * NDERIV, sizeof(double));
  /* common.f90:379-388: vold, uold, vnewstep, unewstep hold krh..kvold_end */
  G.kvold_end = G.mhd ? G.kbz + 1 : 5;   /* up to B^eta, after any viscous variables */
  G.vnewstep = calloc(n * G.kvold_end, sizeof(double));
  G.vold = calloc(n * G.kvold_end, sizeof(double));
  G.unewstep = calloc(n * G.kvold_end, sizeof(double));
  G.uold = calloc(n * G.kvold_end, sizeof(double));
  G.ax = calloc((size_t)(G.nx + 1) * G.ny * G.nz * 2, sizeof(double));
  G.ay = calloc((size_t)G.nx * (G.ny + 1) * G.nz * 2, sizeof(double));
  G.az = calloc((size_t)G.nx * G.ny * (G.nz + 1) * 2, sizeof(double));
  G.t = G.tmin; G.timeold = G.t; G.timeinterval = 0.; G.g_cov_3_old = 0.;
  system_metric();
  G.dt = G.maxdt;
  G.tout = G.t + G.dtout;
  G.iter = 0; G.run_crashed = 0; G.errcode = 0;
  G.smooth_first = 1;
  G.temp_end = 0.; G.check_max = 0.; G.freeze_type = 0; G.temp_check = 0;
  memset(G.nrep, 0, sizeof G.nrep);
  G.crash_cell = -1;
  return G.nv;
}

/* init.f90:2937-2958: per-cell prim2cons followed by cons2prim */
int orc_init_from_prim(const double *vin) {
  size_t n = ncell();
  double vv[MAXNV], uu[MAXNV];
  for (int iz = 0; iz < G.nz; iz++)
    for (int iy = 0; iy < G.ny; iy++)
      for (int ix = 0; ix < G.nx; ix++) {
        size_t c = cell_index(ix, iy, iz);
        for (int k = 0; k < G.nv; k++) vv[k] = vin[k * n + c];
        int e = G.rmhd ? rmhd_prim2cons(vv, uu, 1) : G.viscosity ? sys_prim2cons_0(vv, uu) : sys_prim2cons(vv, uu);
        if (e > 0) return e;
        if (!G.viscosity) {
          e = sys_cons2prim(uu, vv);
          if (e > 0) return e;
        }
        for (int k = 0; k < G.nv; k++) { G.v[k * n + c] = vv[k]; G.u[k * n + c] = uu[k]; }
      }
  return 0;
}

void orc_set_state(const double *v, const double *u) {
  size_t n = ncell() * G.nv;
  memcpy(G.v, v, n * sizeof(double));
  memcpy(G.u, u, n * sizeof(double));
}

void orc_get_state(double *v, double *u) {
  size_t n = ncell() * G.nv;
  memcpy(v, G.v, n * sizeof(double));
  memcpy(u, G.u, n * sizeof(double));
}

void orc_get_deriv(double *d) { memcpy(d, G.deriv, ncell() * NDERIV * sizeof(double)); }

void orc_get_repair_counts(long *out) { memcpy(out, G.nrep, sizeof G.nrep); }

long orc_crash_cell(void) { return G.crash_cell; }

int orc_get_repair_cells(long *out, int n) {
  int m = G.nrep[REP_U2V] < REPLOG ? (int)G.nrep[REP_U2V] : REPLOG;
  if (m > n) m = n;
  memcpy(out, G.replog, sizeof(long) * m);
  return m;
}

void orc_get_scalars(double *s) {
  s[0] = G.t; s[1] = G.dt; s[2] = G.timeold; s[3] = G.timeinterval;
  s[4] = G.g_cov_3_old; s[5] = G.glm_ch_old; s[6] = G.tout; s[7] = G.iter;
  s[8] = G.run_crashed; s[9] = G.errcode;
}

/* ------------------------------------------------------------------ */
/* Boundary fill of one pencil variable, evolve.f90:872-1127 (periodic 0,
 * reflect +-1, zeroth-order outflow 2, third-order extrapolation 3).  p
 * points at the first interior element; ghosts p[-NGC..-1], p[m..m+NGC-1]. */
static void pencil_bc(double *p, int m, int ibc) {
  switch (ibc) {
    case 0:
      for (int l = 1; l <= NGC; l++) { p[-NGC + l - 1] = p[m - NGC + l - 1]; p[m + l - 1] = p[l - 1]; }
      break;
    case 1:
      for (int l = 1; l <= NGC; l++) { p[-NGC + l - 1] = p[NGC - l]; p[m + l - 1] = p[m - l]; }
      break;
    case -1:
      for (int l = 1; l <= NGC; l++) { p[-NGC + l - 1] = -p[NGC - l]; p[m + l - 1] = -p[m - l]; }
      break;
    case 3:
      for (int l = 1; l <= NGC; l++) {
        p[-l] = p[-l + 3] - 3. * (p[-l + 2] - p[-l + 1]);
      }
      for (int l = 1; l <= NGC; l++) {
        p[m + l - 1] = p[m + l - 4] - 3. * (p[m + l - 3] - p[m + l - 2]);
      }
      break;
    default: /* 2 */
      for (int l = 1; l <= NGC; l++) { p[-NGC + l - 1] = p[0]; p[m + l - 1] = p[m - 1]; }
      break;
  }
}

/* HLL of work.f90:195-229 on one interface: from wl/wr primitives to the
 * flux w1[0..nv-1] and speeds w1[nv], w1[nv+1].  Returns an error code. */
static int hll_face(double *wl, double *wr, int dir, double *w1) {
  double ul[MAXNV], ur[MAXNV], fl[MAXNV], fr[MAXNV], vfl[2], vfr[2];
  int e = sys_flux(wl, ul, fl, vfl, dir);
  if (e > 0) return e;
  e = sys_flux(wr, ur, fr, vfr, dir);
  if (e > 0) return e;
  double ap = fmax2(fmax2(0., vfl[0]), vfr[0]);
  double am = fmax2(fmax2(0., -vfl[1]), -vfr[1]);
  G.aflux[dir - 1] = fmax2(ap, am);
  double a1 = (ap == 0.0 && am == 0.0) ? 1. : 1. / (ap + am);
  /* resistive system: B, E and psi travel at the speed of light, whose
   * coordinate value is 1/sqrt(g_dir); HLL with those speeds for them */
  const int kem = G.rmhd ? G.kbx : G.nv;
  for (int k = 0; k < kem; k++) w1[k] = a1 * (ap * fl[k] + am * fr[k] - ap * am * (ur[k] - ul[k]));
  if (G.rmhd) {
    const double c = 1. / sqrt(G.g_cov[dir - 1]);
    for (int k = kem; k < G.nv; k++) w1[k] = 0.5 * (fl[k] + fr[k] - c * (ur[k] - ul[k]));
  }
  w1[G.nv] = ap;
  w1[G.nv + 1] = am;
  return 0;
}

/* One directional sweep, evolve_df{x,y,z} (evolve.f90:302-685).  dir = 1,2,3.
 * Pencils are gathered along dir with NGC ghosts filled by pencil_bc; the
 * Lorentz-velocity reconstruction feeds deriv only for viscous runs. */
static int sweep(int dir) {
  const int nx = G.nx, ny = G.ny, nz = G.nz, nv = G.nv;
  const int m = dir == 1 ? nx : dir == 2 ? ny : nz;
  const int n1 = dir == 1 ? ny : nx;           /* first transverse index */
  const int n2 = dir == 3 ? ny : nz;           /* second transverse index */
  const double dd = dir == 1 ? G.ddx : dir == 2 ? G.ddy : G.ddz;
  const int len = m + 2 * NGC;
  double *w0 = malloc(sizeof(double) * len * nv);
  double *wd = malloc(sizeof(double) * len * 4);
  double *wl = malloc(sizeof(double) * (m + 1) * nv), *wr = malloc(sizeof(double) * (m + 1) * nv);
  double *wdl = malloc(sizeof(double) * (m + 1) * 4), *wdr = malloc(sizeof(double) * (m + 1) * 4);
  double *w1 = malloc(sizeof(double) * (m + 1) * (nv + 2));
  const size_t nc = ncell();
  int error_flag = 0;
  const double *g = G.g_cov;
  double fl[MAXNV], fr[MAXNV];
  for (int j2 = 0; j2 < n2 && !error_flag; j2++) {
    for (int j1 = 0; j1 < n1; j1++) {
      int ix = 0, iy = 0, iz = 0;
      /* gather the pencil */
      for (int i = 0; i < m; i++) {
        if (dir == 1) { ix = i; iy = j1; iz = j2; }
        else if (dir == 2) { ix = j1; iy = i; iz = j2; }
        else { ix = j1; iy = j2; iz = i; }
        size_t c = cell_index(ix, iy, iz);
        for (int k = 0; k < nv; k++) w0[k * len + NGC + i] = G.v[k * nc + c];
      }
      for (int k = 0; k < nv; k++) pencil_bc(&w0[k * len + NGC], m, G.boundary_cond);
      if (G.viscosity) {
        for (int i = 0; i < len; i++) {
          double vx = w0[KVX * len + i], vy = w0[KVY * len + i], vz = w0[KVZ * len + i];
          double gl = 1. / sqrt(1. - vx * vx - vy * vy - g[2] * (vz * vz));
          wd[0 * len + i] = gl;
          wd[1 * len + i] = gl * vx;
          wd[2 * len + i] = gl * vy;
          wd[3 * len + i] = gl * vz;
        }
      }
      for (int k = 0; k < nv; k++)
        rec_pencil(G.recal, &w0[k * len + NGC - 1], &wl[k * (m + 1)], &wr[k * (m + 1)], m);
      if (G.viscosity)
        for (int k = 0; k < 4; k++)
          rec_pencil(G.recal, &wd[k * len + NGC - 1], &wdl[k * (m + 1)], &wdr[k * (m + 1)], m);
      /* reconstruction repair, evolve.f90:334-369.  Index i is the interface
       * i+1/2; the reference repairs both sides with the cell left of it
       * (w0(ix) with pencil index ix = i, i.e. cell i of the 1-based pencil,
       * which is w0[NGC + i - 1] here). */
      for (int i = 0; i <= m; i++) {
        double *c0 = &w0[NGC + i - 1];
#define W0(k) c0[(k) * len]
#define WR(k) wr[(k) * (m + 1) + i]
#define WL(k) wl[(k) * (m + 1) + i]
        if (WR(KPR) <= 0) { WR(KPR) = W0(KPR); G.nrep[REP_REC_PR]++; }
        double v2 = WR(KVX) * WR(KVX) * g[0] + WR(KVY) * WR(KVY) * g[1] + WR(KVZ) * WR(KVZ) * g[2];
        if (v2 >= 1.) {
          G.nrep[REP_REC_V]++;
          WR(KVX) = W0(KVX); WR(KVY) = W0(KVY); WR(KVZ) = W0(KVZ);
          v2 = WR(KVX) * WR(KVX) * g[0] + WR(KVY) * WR(KVY) * g[1] + WR(KVZ) * WR(KVZ) * g[2];
          if (v2 >= 1) {
            G.nrep[REP_REC_VMAX]++;
            double s = sqrt(v2);
            for (int k = KVX; k <= KVZ; k++) { W0(k) = W0(k) * G.maxspeed / s; WR(k) = W0(k); }
          }
        }
        if (WL(KPR) <= 0) { WL(KPR) = W0(KPR); G.nrep[REP_REC_PR]++; }
        v2 = WL(KVX) * WL(KVX) * g[0] + WL(KVY) * WL(KVY) * g[1] + WL(KVZ) * WL(KVZ) * g[2];
        if (v2 >= 1.) {
          G.nrep[REP_REC_V]++;
          WL(KVX) = W0(KVX); WL(KVY) = W0(KVY); WL(KVZ) = W0(KVZ);
          v2 = WL(KVX) * WL(KVX) * g[0] + WL(KVY) * WL(KVY) * g[1] + WL(KVZ) * WL(KVZ) * g[2];
          if (v2 >= 1) {
            G.nrep[REP_REC_VMAX]++;
            double s = sqrt(v2);
            for (int k = KVX; k <= KVZ; k++) { W0(k) = W0(k) * G.maxspeed / s; WL(k) = W0(k); }
          }
        }
#undef W0
#undef WR
#undef WL
      }
      /* deriv(d?x) from face-averaged Lorentz velocities, evolve.f90:371-374 */
      if (G.viscosity) {
        static const int slot[3][4] = {{DTX, DXX, DYX, DZX}, {DTY, DXY, DYY, DZY}, {DTZ, DXZ, DYZ, DZZ}};
        for (int i = 1; i <= m; i++) {
          if (dir == 1) { ix = i - 1; iy = j1; iz = j2; }
          else if (dir == 2) { ix = j1; iy = i - 1; iz = j2; }
          else { ix = j1; iy = j2; iz = i - 1; }
          for (int q = 0; q < 4; q++) {
            double r = dd * 0.5 * (wdl[q * (m + 1) + i] + wdr[q * (m + 1) + i]
                                   - wdl[q * (m + 1) + i - 1] - wdr[q * (m + 1) + i - 1]);
            DERIV(ix, iy, iz, slot[dir - 1][q]) = r;
          }
        }
      }
      /* work_glm{x,y,z}, work.f90:67-139: Dedner interface values of psi and
       * the normal B.  The reference sets wr(psi) = wl(psi) before using
       * their difference, so the B update reduces to the average. */
      if (G.divclean) {
        const int kb = dir == 1 ? G.kbx : dir == 2 ? G.kby : G.kbz, kg = G.kglm;
        for (int i = 0; i <= m; i++) {
          double *gl = &wl[kg * (m + 1) + i], *gr = &wr[kg * (m + 1) + i];
          double *bl = &wl[kb * (m + 1) + i], *br = &wr[kb * (m + 1) + i];
          *gl = 0.5 * (*gl + *gr) - (*br - *bl) * G.glm_ch / 2;
          *gr = *gl;
          *bl = 0.5 * (*bl + *br) - (*gr - *gl) / (2 * G.glm_ch);
          *br = *bl;
        }
      }
      /* work_fup? HLL, work.f90:195-229 */
      for (int i = 0; i <= m; i++) {
        for (int k = 0; k < nv; k++) { fl[k] = wl[k * (m + 1) + i]; fr[k] = wr[k * (m + 1) + i]; }
        int e = hll_face(fl, fr, dir, &w1[i * (nv + 2)]);
        if (e > 0) { error_flag = e; break; }
      }
      if (error_flag) break;
      /* du += dd*(F(i)-F(i-1)), evolve.f90:383; ax = (ap,am), 387 */
      for (int i = 1; i <= m; i++) {
        if (dir == 1) { ix = i - 1; iy = j1; iz = j2; }
        else if (dir == 2) { ix = j1; iy = i - 1; iz = j2; }
        else { ix = j1; iy = j2; iz = i - 1; }
        size_t c = cell_index(ix, iy, iz);
        for (int k = 0; k < G.nu; k++)
          G.du[k * nc + c] = G.du[k * nc + c] + dd * (w1[i * (nv + 2) + k] - w1[(i - 1) * (nv + 2) + k]);
      }
      for (int i = 0; i <= m; i++) {
        double *a;
        if (dir == 1) a = &G.ax[(((size_t)j2 * ny + j1) * (nx + 1) + i) * 2];
        else if (dir == 2) a = &G.ay[(((size_t)j2 * (ny + 1) + i) * nx + j1) * 2];
        else a = &G.az[(((size_t)i * ny + j2) * nx + j1) * 2];
        a[0] = w1[i * (nv + 2) + nv];
        a[1] = w1[i * (nv + 2) + nv + 1];
      }
    }
  }
  free(w0); free(wd); free(wl); free(wr); free(wdl); free(wdr); free(w1);
  if (error_flag) { G.errcode = error_flag; G.run_crashed = 1; }
  return error_flag;
}

/* dts, viscous.f90:56-97: lagged first-order time derivatives of u^mu */
static void dts(void) {
  size_t nc = ncell();
  if (G.timeinterval == 0) {
    for (int k = DTT; k <= DZT; k++) memset(&G.deriv[k * nc], 0, nc * sizeof(double));
    return;
  }
  const double *g = G.g_cov;
  for (size_t c = 0; c < nc; c++) {
    const double *qo = G.vold, *qn = G.vnewstep;
    double ox = qo[KVX * nc + c], oy = qo[KVY * nc + c], oz = qo[KVZ * nc + c];
    double nxv = qn[KVX * nc + c], nyv = qn[KVY * nc + c], nzv = qn[KVZ * nc + c];
    double gamma1 = 1. / sqrt(1. - g[0] * (ox * ox) - g[1] * (oy * oy) - G.g_cov_3_old * (oz * oz));
    double gamma2 = 1. / sqrt(1. - g[0] * (nxv * nxv) - g[1] * (nyv * nyv) - g[2] * (nzv * nzv));
    G.deriv[DTT * nc + c] = (gamma2 - gamma1) / G.timeinterval;
    G.deriv[DXT * nc + c] = (gamma2 * nxv - gamma1 * ox) / G.timeinterval;
    G.deriv[DYT * nc + c] = (gamma2 * nyv - gamma1 * oy) / G.timeinterval;
    G.deriv[DZT * nc + c] = (gamma2 * nzv - gamma1 * oz) / G.timeinterval;
  }
}

/* evolve_src, evolve.f90:689-718 */
static int evolve_src(void) {
  size_t nc = ncell();
  if (G.viscosity) dts();
  double vv[MAXNV], src[MAXNV], srcs[MAXNV];
  for (int iz = 0; iz < G.nz; iz++)
    for (int iy = 0; iy < G.ny; iy++)
      for (int ix = 0; ix < G.nx; ix++) {
        size_t c = cell_index(ix, iy, iz);
        for (int k = 0; k < G.nv; k++) vv[k] = G.v[k * nc + c];
        int e = sys_source(ix, iy, iz, vv, src, srcs);
        if (e > 0) { G.errcode = e; return e; }
        for (int k = 0; k < G.nu; k++) {
          G.du[k * nc + c] = G.du[k * nc + c] - src[k];
          G.du_stiff[k * nc + c] = G.du_stiff[k * nc + c] - srcs[k];
        }
      }
  return 0;
}

/* evolve_glm_source, evolve.f90:83-110 */
static void evolve_glm_source(double dt_glm) {
  double expfact = exp(-(G.glm_ch * G.glm_ch) * G.glm_alpha * dt_glm);
  size_t nc = ncell();
  double *a = &G.u[(size_t)G.kglm * nc];
  for (size_t c = 0; c < nc; c++) a[c] = a[c] * expfact;
}

/* evolve_dt, evolve.f90:723-764 */
static double evolve_dt(void) {
  if (G.rmhd) {
    /* resistive system: the speed of light in each active direction */
    double amax = 0.;
    if (G.nx > 1) amax = fmax2(amax, G.ddx);
    if (G.ny > 1) amax = fmax2(amax, G.ddy);
    if (G.nz > 1) amax = fmax2(amax, G.ddz / sqrt(G.g_cov[2]));
    return G.cfl / amax;
  }
  if (G.divclean) return G.cfl / 1.;
  double amax = 0.;
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  for (int iz = 0; iz < nz; iz++)
    for (int iy = 0; iy < ny; iy++)
      for (int ix = 0; ix < nx; ix++) {
        if (nx > 1) {
          const double *a = &G.ax[(((size_t)iz * ny + iy) * (nx + 1) + ix) * 2];
          double m = fmax2(fmax2(fmax2(a[0], a[1]), a[2]), a[3]);
          amax = fmax2(amax, G.ddx * m);
        }
        if (ny > 1) {
          const double *a0 = &G.ay[(((size_t)iz * (ny + 1) + iy) * nx + ix) * 2];
          const double *a1 = &G.ay[(((size_t)iz * (ny + 1) + iy + 1) * nx + ix) * 2];
          double m = fmax2(fmax2(fmax2(a0[0], a1[0]), a0[1]), a1[1]);
          amax = fmax2(amax, G.ddy * m);
        }
        if (nz > 1) {
          const double *a0 = &G.az[(((size_t)iz * ny + iy) * nx + ix) * 2];
          const double *a1 = &G.az[(((size_t)(iz + 1) * ny + iy) * nx + ix) * 2];
          double m = fmax2(fmax2(fmax2(a0[0], a1[0]), a0[1]), a1[1]);
          amax = fmax2(amax, G.ddz * m);
        }
      }
  if (amax == 0) {
    if (nx > 1) amax = fmax2(amax, G.ddx);
    if (ny > 1) amax = fmax2(amax, G.ddy);
    if (nz > 1) amax = fmax2(amax, G.ddz);
  }
  return G.cfl / amax;
}

/* evolve_u2v, evolve.f90:769-868.  The neighbour-average repair reads the
 * six face neighbours; on boundary cells the reference reads outside the
 * array (its u has no ghosts), here the index is clamped to the grid. */
static int evolve_u2v(void) {
  const size_t nc = ncell();
  const double prlimit = 0.01;
  (void)prlimit;
  double uu[MAXNV], vv[MAXNV];
  for (int iz = 0; iz < G.nz; iz++)
    for (int iy = 0; iy < G.ny; iy++)
      for (int ix = 0; ix < G.nx; ix++) {
        size_t c = cell_index(ix, iy, iz);
        for (int k = 0; k < G.nv; k++) { vv[k] = G.v[k * nc + c]; uu[k] = G.u[k * nc + c]; }
        /* inside an IMEX stage of the resistive system the inversion also
         * solves the stiff relation for E (and changes u there) */
        int e = (G.rmhd && G.imex_adt > 0.) ? rmhd_cons2prim_implicit(uu, vv, G.imex_adt) : sys_cons2prim(uu, vv);
        if (e != 0) {
          if (G.nrep[REP_U2V] < REPLOG) G.replog[G.nrep[REP_U2V]] = (long)c;
          G.nrep[REP_U2V]++;
          int xl = ix > 0 ? ix - 1 : ix, xr = ix < G.nx - 1 ? ix + 1 : ix;
          int yl = iy > 0 ? iy - 1 : iy, yr = iy < G.ny - 1 ? iy + 1 : iy;
          int zl = iz > 0 ? iz - 1 : iz, zr = iz < G.nz - 1 ? iz + 1 : iz;
          size_t n6[6] = {cell_index(xl, iy, iz), cell_index(xr, iy, iz), cell_index(ix, yl, iz),
                          cell_index(ix, yr, iz), cell_index(ix, iy, zl), cell_index(ix, iy, zr)};
          for (int k = 0; k < G.nu; k++) {
            double *a = &G.u[k * nc];
            a[c] = (a[n6[0]] + a[n6[1]] + a[n6[2]] + a[n6[3]] + a[n6[4]] + a[n6[5]]) / 6.;
            uu[k] = a[c];
          }
          sys_cons2prim(uu, vv);
          const double *g = G.g_cov;
          double v2 = vv[KVX] * vv[KVX] + vv[KVY] * vv[KVY] + vv[KVZ] * vv[KVZ] * g[2];
          if (v2 > 1) {
            G.nrep[REP_U2V_V]++;
            double s = G.maxspeed / sqrt(v2);
            vv[KVX] = s * vv[KVX]; vv[KVY] = s * vv[KVY]; vv[KVZ] = s * vv[KVZ];
          }
          if (vv[KPR] < 0) {
            G.nrep[REP_U2V_PR]++;
            const double *p = &G.v[KPR * nc];
            vv[KPR] = (p[n6[0]] + p[n6[1]] + p[n6[2]] + p[n6[3]] + p[n6[4]] + p[n6[5]]) / 6.;
          }
          sys_prim2cons(vv, uu);
          sys_cons2prim(uu, vv);
          v2 = vv[KVX] * vv[KVX] + vv[KVY] * vv[KVY] + vv[KVZ] * vv[KVZ] * g[2];
          e = ((vv[KPR] > 0) && (v2 < 1)) ? 0 : 1;
          if (e) { G.nrep[REP_U2V_FAIL]++; G.crash_cell = (long)c; }
        }
        if (G.viscosity && G.obtained_zz) G.v[KPIZZ * nc + c] = vv[KPIZZ];
        /* u is written before the error check (evolve.f90:848-849): a cell
         * whose repair failed keeps its last attempt in u */
        for (int k = 0; k < G.nu; k++) G.u[k * nc + c] = uu[k];
        if (e > 0) { G.errcode = e; G.run_crashed = 1; return e; }
        for (int k = 0; k < G.nv; k++) { G.v[k * nc + c] = vv[k]; G.u[k * nc + c] = uu[k]; }
      }
  return 0;
}

/* evolve_smooth_viscosity, evolve.f90:1159-1207.  The threshold pressure
 * comes from find_pressure_from_temperature (eos.f90:1221-1260), restated
 * here as a bisection on T(p) over the EoS range. */
static int evolve_smooth_viscosity(void) {
  const size_t nc = ncell();
  if (G.smooth_first) {
    find_pressure_from_temperature(G.smooth_temp, &G.smooth_pressure);
    G.smooth_first = 0;
  }
  double vv[MAXNV], uu[MAXNV];
  for (size_t c = 0; c < nc; c++) {
    double pr = G.v[KPR * nc + c];
    if (pr < G.smooth_pressure) {
      double r = pr / G.smooth_pressure;
      for (int k = 0; k < G.nv; k++) vv[k] = G.v[k * nc + c];
      for (int k = KPIBU; k <= KPIZZ; k++) vv[k] = vv[k] * (r * r);
      if (G.obtained_zz) {
        double a, b, cc, d, pizz;
        get_derived_pi_zz(vv, &a, &b, &cc, &d, &pizz);
        vv[KPIZZ] = pizz;
      }
      int e = sys_prim2cons(vv, uu);
      if (e > 0) { G.errcode = e; G.run_crashed = 1; return e; }
      for (int k = 0; k < G.nv; k++) { G.v[k * nc + c] = vv[k]; G.u[k * nc + c] = uu[k]; }
    }
  }
  return 0;
}

/* find_pressure_from_temperature, eos.f90:1221-1252.  The reference leaves
 * the lower bisection bound pr_low uninitialised; it starts at 0 here.  The
 * bisection stops at |T(p) - T_cut| <= 1e-10 T_cut, so the threshold agrees
 * with the reference to that tolerance whatever the starting bound. */
int find_pressure_from_temperature(double temp_input, double *pres) {
  const double tolerance = 1.e-10;
  double en_try = G.eqstate == 3 ? G.tab_e[0] : 0.;
  if (en_try == 0.) en_try = tolerance / 2.;
  double pr_try, d1, d2, en, temp_try, s;
  eos_pressure(1., en_try, &pr_try, &d1, &d2);
  get_derived_data(1., pr_try, &en, &temp_try, &s);
  while (temp_try <= temp_input) {
    get_derived_data(1., pr_try, &en, &temp_try, &s);
    pr_try = pr_try * 1.1;
  }
  double pr_low = 0., pr_high = pr_try;
  while (fabs(temp_try - temp_input) > temp_input * tolerance) {
    pr_try = (pr_low + pr_high) / 2.;
    get_derived_data(1., pr_try, &en, &temp_try, &s);
    if (temp_try >= temp_input) pr_high = pr_try;
    else pr_low = pr_try;
  }
  *pres = pr_try;
  return 0;
}

/* evolve_main for explicit RK2/RK3, evolve.f90:114-203 + 276-295 */
/* the fluxes and sources of one stage into du / du_stiff slot irk */
static int ssp_rhs(int irk) {
  const size_t n = ncell() * G.nu;
  G.du = G.du_all + (size_t)(irk - 1) * n;
  G.du_stiff = G.dus_all + (size_t)(irk - 1) * n;
  memset(G.du, 0, n * sizeof(double));
  memset(G.du_stiff, 0, n * sizeof(double));
  G.glm_ch = G.glm_ch_old;
  if (G.nx > 1 && sweep(1)) return G.errcode;
  if (G.ny > 1 && sweep(2)) return G.errcode;
  if (G.nz > 1 && sweep(3)) return G.errcode;
  if (evolve_src()) { G.run_crashed = 1; return G.errcode; }
  G.glm_ch_old = G.rmhd ? 1. : fmax2(fmax2(G.aflux[0], G.aflux[1]), G.aflux[2]);
  return 0;
}

/* the stiff rate of the first IMEX stage of the resistive system: the
 * explicit rate of K1 has weight zero in every later stage, so only
 * du_stiff = -src_stiff is formed (cell by cell, no sweep) */
static void ssp_stiff_only(void) {
  const size_t nc = ncell(), n = nc * G.nu;
  G.du = G.du_all;
  G.du_stiff = G.dus_all;
  memset(G.du, 0, n * sizeof(double));
  memset(G.du_stiff, 0, n * sizeof(double));
  double vv[MAXNV], s3[3];
  for (size_t c = 0; c < nc; c++) {
    for (int k = 0; k < G.nv; k++) vv[k] = G.v[k * nc + c];
    rmhd_stiff(vv, s3);
    for (int k = 0; k < 3; k++) G.du_stiff[(size_t)(G.kex + k) * nc + c] = -s3[k];
  }
}

/* evolve_main for the IMEX-SSP3(4,3,3) scheme (nrk = SSP),
 * evolve.f90:127-130, 204-264: explicit stages from u0 and the stored
 * du of stages 2-4, stiff parts du_stiff of stages 1-4.  For ideal MHD the
 * stiff part is zero and the stages are the reference's; for the resistive
 * system every stage solves its implicit part inside evolve_u2v, the first
 * one included (the reference evaluates the stage-1 stiff part at u^n).
 * After the last stage the final combination and evolve_u2v. */
static int evolve_main_ssp(int irk) {
  const double alpha = 0.24169426078821, beta = 0.06042356519705, eta = 0.12915286960590;
  const double delta = 0.5 - beta - eta - alpha;
  const size_t n = ncell() * G.nu;
  double *du1 = G.du_all, *du2 = G.du_all + n, *du3 = G.du_all + 2 * n, *du4 = G.du_all + 3 * n;
  double *ds1 = G.dus_all, *ds2 = G.dus_all + n, *ds3 = G.dus_all + 2 * n, *ds4 = G.dus_all + 3 * n;
  (void)du1;
  const double dt = G.dt;
  if (irk == 1) {
    double dt_local = evolve_dt();
    dt_local = fmin2(dt_local, G.maxdt);
    dt_local = fmin2(dt_local, G.tout - G.t);
    if (fabs(G.t + G.dt - G.tout) < 1.e-9) dt_local = G.tout - G.t;
    else if ((G.t + 2. * dt_local - G.tout) > 1.e-9) dt_local = (G.tout - G.t) / 2.;
    G.dt = dt_local;
    memcpy(G.u0, G.u, n * sizeof(double));
    if (G.rmhd) {
      /* the implicit first stage K1 = u^n + alpha dt R(K1) of the scheme,
       * which the reference replaces by K1 = u^n (no inversion at stage 1);
       * with a stiff current that shortcut is not stable */
      G.imex_adt = alpha * G.dt;
      int e = evolve_u2v();
      G.imex_adt = 0.;
      if (e) return e;
      ssp_stiff_only();
    } else {
      if (ssp_rhs(1)) return G.errcode;
    }
    if (G.divclean) evolve_glm_source(G.dt / 2.);
  } else if (irk == 2) {
    for (size_t i = 0; i < n; i++) G.u[i] = G.u0[i] + dt * alpha * ds1[i];
    G.imex_adt = alpha * dt;
    int e = evolve_u2v();
    G.imex_adt = 0.;
    if (e) return e;
    if (ssp_rhs(2)) return G.errcode;
  } else if (irk == 3) {
    for (size_t i = 0; i < n; i++) G.u[i] = G.u0[i] - dt * (du2[i] + (1. - alpha) * ds2[i]);
    G.imex_adt = alpha * dt;
    int e = evolve_u2v();
    G.imex_adt = 0.;
    if (e) return e;
    if (ssp_rhs(3)) return G.errcode;
  } else {
    for (size_t i = 0; i < n; i++)
      G.u[i] = G.u0[i] - dt * (du2[i] / 4. + du3[i] / 4. + beta * ds1[i] + eta * ds2[i] + delta * ds3[i]);
    G.imex_adt = alpha * dt;
    int e = evolve_u2v();
    G.imex_adt = 0.;
    if (e) return e;
    if (ssp_rhs(4)) return G.errcode;
    for (size_t i = 0; i < n; i++)
      G.u[i] = G.u0[i] - dt * (du2[i] / 6. + du3[i] / 6. + du4[i] * 2. / 3. + ds2[i] / 6. + ds3[i] / 6. +
                               ds4[i] * 2. / 3.);
    if (G.divclean) evolve_glm_source(G.dt / 2.);
    if (evolve_u2v()) return G.errcode;
  }
  /* the resistive system smooths the step's result only, not the states
   * of its implicit stages */
  if (G.rmhd && irk < 4) return 0;
  if (G.viscosity && G.enable_smooth_viscosity) {
    if (evolve_smooth_viscosity()) return G.errcode;
  }
  return 0;
}

static int evolve_main(int irk) {
  if (G.nrk == SSP) return evolve_main_ssp(irk);
  static const double crk[2][3][3] = {
    /* [c][irk][nrk]: crk(c,irk,nrk) of evolve.f90:122-125 */
    {{0., 0., 0.}, {0., 1. / 2., 3. / 4.}, {0., 0., 1. / 3.}},
    {{1., 1., 1.}, {0., 1. / 2., 1. / 4.}, {0., 0., 2. / 3.}}};
  const int nrk = G.nrk;
  const double c1 = crk[0][irk - 1][nrk - 1], c2 = crk[1][irk - 1][nrk - 1];
  const size_t n = ncell() * G.nu;
  G.du = G.du_all;
  G.du_stiff = G.dus_all;
  memset(G.du, 0, n * sizeof(double));
  memset(G.du_stiff, 0, n * sizeof(double));
  G.glm_ch = G.glm_ch_old;
  if (G.nx > 1 && sweep(1)) return G.errcode;
  if (G.ny > 1 && sweep(2)) return G.errcode;
  if (G.nz > 1 && sweep(3)) return G.errcode;
  if (evolve_src()) { G.run_crashed = 1; return G.errcode; }
  G.glm_ch_old = G.rmhd ? 1. : fmax2(fmax2(G.aflux[0], G.aflux[1]), G.aflux[2]);
  if (irk == 1) {
    double dt_local = evolve_dt();
    dt_local = fmin2(dt_local, G.maxdt);
    dt_local = fmin2(dt_local, G.tout - G.t);
    if (fabs(G.t + G.dt - G.tout) < 1.e-9) dt_local = G.tout - G.t;
    else if ((G.t + 2. * dt_local - G.tout) > 1.e-9) dt_local = (G.tout - G.t) / 2.;
    G.dt = dt_local;
    if (G.divclean) evolve_glm_source(G.dt / 2.);
    memcpy(G.u0, G.u, n * sizeof(double));
  }
  if (G.rmhd) {
    /* resistive system with an explicit RK: the stiff current explicitly */
    for (size_t i = 0; i < n; i++) G.u[i] = c1 * G.u0[i] + c2 * (G.u[i] - G.dt * (G.du[i] + G.du_stiff[i]));
  } else {
    for (size_t i = 0; i < n; i++) G.u[i] = c1 * G.u0[i] + c2 * (G.u[i] - G.dt * G.du[i]);
  }
  if (G.divclean && irk == nrk) evolve_glm_source(G.dt / 2.);
  if (evolve_u2v()) return G.errcode;
  if (G.viscosity && G.enable_smooth_viscosity) {
    if (evolve_smooth_viscosity()) return G.errcode;
  }
  return 0;
}

/* main-loop body of echo.f90:185-262 without output */
/* check_temp, out.f90:1487-1557: the maximum over the cells of the
 * temperature (freeze_type 0) or energy density (1) from
 * get_derived_temp_for_check (eos.f90:706-759, the interpolation of
 * get_derived_data); temp_check = 1 below temp_end. */
static void check_temp(void) {
  double mx = 0.;
  for (int iz = 0; iz < G.nz; iz++)
    for (int iy = 0; iy < G.ny; iy++)
      for (int ix = 0; ix < G.nx; ix++) {
        double en = 0., temp = 0., s;
        get_derived_data(V(ix, iy, iz, KRH), V(ix, iy, iz, KPR), &en, &temp, &s);
        mx = fmax2(mx, G.freeze_type == 0 ? temp : en);
      }
  G.check_max = mx;
  G.temp_check = mx < G.temp_end ? 1 : 0;
}

void orc_set_stop(double temp_end, int freeze_type) {
  G.temp_end = temp_end;
  G.freeze_type = freeze_type;
  G.temp_check = 0;
  G.check_max = 0.;
}

void orc_get_stop(int *stopped, double *max_value) {
  *stopped = G.temp_check;
  *max_value = G.check_max;
}

int orc_steps(int nsteps) {
  const size_t nc = ncell();
  for (int s = 0; s < nsteps; s++) {
    /* echo.f90:189-190: the check opens each pass of the time loop */
    if (G.temp_check) break;
    if (G.temp_end > 0) {
      check_temp();
      if (G.temp_check) break;
    }
    memcpy(G.vnewstep, G.v, G.kvold_end * nc * sizeof(double));
    memcpy(G.unewstep, G.u, G.kvold_end * nc * sizeof(double));
    for (int irk = 1; irk <= G.nrk; irk++) {
      int e = evolve_main(irk);
      if (e || G.run_crashed) { G.run_crashed = 1; return e ? e : 1; }
    }
    G.timeold = G.t;
    G.t = G.t + G.dt;
    G.g_cov_3_old = G.g_cov[2];
    system_metric();
    G.iter++;
    G.timeinterval = G.dt;
    memcpy(G.vold, G.vnewstep, G.kvold_end * nc * sizeof(double));
    memcpy(G.uold, G.unewstep, G.kvold_end * nc * sizeof(double));
    if ((G.t > G.tout) || ((G.tout - G.t) < 1.e-10)) {
      G.tout = G.tout + G.dtout;
      if ((fabs(G.tout - G.tmax) < 1.e-10) || (G.tout > G.tmax)) G.tout = G.tmax;
    }
  }
  return 0;
}

void orc_set_maxdt(double maxdt) { G.maxdt = maxdt; }

/* evolve_u2v alone on the current v, u (host-forced repair states) */
int orc_u2v(void) {
  G.run_crashed = 0; G.errcode = 0;
  int e = evolve_u2v();
  return e ? e : (G.run_crashed ? 1 : 0);
}

/* one RK stage (evolve_main) without the step bookkeeping of orc_steps;
 * irk = 1 also takes the step-start snapshot */
int orc_stage(int irk) {
  const size_t nc = ncell();
  if (irk == 1) {
    memcpy(G.vnewstep, G.v, G.kvold_end * nc * sizeof(double));
    memcpy(G.unewstep, G.u, G.kvold_end * nc * sizeof(double));
  }
  int e = evolve_main(irk);
  if (e || G.run_crashed) { G.run_crashed = 1; return e ? e : 1; }
  return 0;
}

/* ------------------------------------------------------------------ */
/* pointwise wrappers */
void orc_set_time(double t) { G.t = t; system_metric(); }

void orc_rec(int code, int n, int ngc_in, const double *f, double *gl, double *gr) {
  rec_pencil(code, f + ngc_in - 1, gl, gr, n);
}
int orc_prim2cons(double *v, double *u) { return sys_prim2cons(v, u); }
int orc_cons2prim(double *u, double *v) { return sys_cons2prim(u, v); }
int orc_flux(double *v, int dir, double *u, double *f, double *vf) { return sys_flux(v, u, f, vf, dir); }
int orc_eos_pressure(double rh, double en, double *o) { return eos_pressure(rh, en, &o[0], &o[1], &o[2]); }
int orc_eos_energy(double rh, double pr, double *en) { return eos_energy(rh, pr, en); }
int orc_eos_sound(double rh, double en, double pr, double *cs2) { return eos_sound(rh, en, pr, cs2); }
int orc_eos_derived(double rh, double pr, double *o) { return get_derived_data(rh, pr, &o[0], &o[1], &o[2]); }
int orc_eos_table(int n, double *t, double *e, double *p, double *c) {
  int m = n < G.entries ? n : G.entries;
  for (int i = 0; i < m; i++) { t[i] = G.tab_t[i]; e[i] = G.tab_e[i]; p[i] = G.tab_p[i]; c[i] = G.tab_c[i]; }
  return G.entries;
}
double orc_przero(void) { return G.przero; }
