/* ORACLE - test infrastructure, not product code (see oracle.h).
 * Pointwise physics: restates system.f90 and viscous.f90 (reference) for
 * ideal and second-order (Israel-Stewart) viscous hydrodynamics in
 * Minkowski or Bjorken (Milne) coordinates.
 */
#include <stdio.h>
#include "oracle_int.h"

static const double HBAR = 0.197326;   /* common.f90:47 */

/* system_metric_minkowski / _bjorken, system.f90:132-160 */
void system_metric(void) {
  G.g_cov[0] = 1.; G.g_cov[1] = 1.;
  if (G.coordinates == BJORKEN) {
    G.g_cov[2] = G.t * G.t;
    G.gp = G.t;
    G.gm = 1. / G.gp;
  } else {
    G.g_cov[2] = 1.; G.gp = 1.; G.gm = 1.;
  }
  G.g_cov0 = -1.;
}

/* get_derived_pi, viscous.f90:227-239 */
void get_derived_pi(const double *p, double *pitt, double *pitx, double *pity, double *pitz) {
  const double *g = G.g_cov;
  *pitx = p[KPIXX] * p[KVX] * g[0] + p[KPIXY] * p[KVY] * g[1] + p[KPIXZ] * p[KVZ] * g[2];
  *pity = p[KPIXY] * p[KVX] * g[0] + p[KPIYY] * p[KVY] * g[1] + p[KPIYZ] * p[KVZ] * g[2];
  *pitz = p[KPIXZ] * p[KVX] * g[0] + p[KPIYZ] * p[KVY] * g[1] + p[KPIZZ] * p[KVZ] * g[2];
  *pitt = *pitx * p[KVX] * g[0] + *pity * p[KVY] * g[1] + *pitz * p[KVZ] * g[2];
}

/* get_derived_pi_zz, viscous.f90:243-279: pi^tt, pi^tx, pi^ty, pi^tz from
 * orthogonality and pi^zz from tracelessness */
void get_derived_pi_zz(const double *p, double *pitt, double *pitx, double *pity, double *pitz, double *pizz) {
  const double g1 = G.g_cov[0], g2 = G.g_cov[1], g3 = G.g_cov[2];
  const double vx = p[KVX], vy = p[KVY], vz = p[KVZ];
  const double xx = p[KPIXX], yy = p[KPIYY], xy = p[KPIXY], xz = p[KPIXZ], yz = p[KPIYZ];
  *pitt = 1. / (1. - vz * vz * g3) * (vx * g1 * xy * vy * g2 + vx * vx * g1 * g1 * xx +
          vy * g1 * yz * vz * g3 + 2. * vx * g1 * xz * vz * g3 + vy * g1 * g1 * xy *
          vx + vy * vy * g1 * yy * g2 + vz * g3 * yz * vy * g2 - xx * g1 * vz * vz *
          g3 - yy * g2 * vz * vz * g3);
  *pitx = xx * vx * g1 + xy * vy * g2 + xz * vz * g3;
  *pity = xy * vx * g1 + yy * vy * g2 + yz * vz * g3;
  *pitz = 1. / (1. - vz * vz * g3) * (xz * vx * g1 + xz * vx * g1 * vz * vz * g3 + yz * vy *
          g2 + vz * vx * g1 * xy * vy * g2 - vz * xx * g1 - vz * yy * g2 + vz * vx *
          vx * g1 * g1 * xx + vy * g1 * yz * vz * vz * g3 + vz * vy * g1 * g1 *
          xy * vx + vz * vy * vy * g1 * yy * g2);
  *pizz = 1. / (g3 * (1. - vz * vz * g3)) * (vx * g1 * xy * vy * g2 - xx * g1 - yy * g2 +
          vx * vx * g1 * g1 * xx + vy * g1 * yz * vz * g3 + 2. * vx * g1 * xz * vz *
          g3 + vy * g1 * g1 * xy * vx + vy * vy * g1 * yy * g2 + vz * g3 * yz *
          vy * g2);
}

/* the pi^{0i} helper as selected by TRACE_IMP (system.f90:226-232) */
static void derived_pi(double *v, double *pitt, double *pitx, double *pity, double *pitz) {
  if (!G.obtained_zz) {
    get_derived_pi(v, pitt, pitx, pity, pitz);
  } else {
    double pizz;
    get_derived_pi_zz(v, pitt, pitx, pity, pitz, &pizz);
    v[KPIZZ] = pizz;
  }
}

/* system_prim2cons, system.f90:164-263 (HD and viscous branches) */
int sys_prim2cons(double *v, double *u) {
  double rh = v[KRH], vx = v[KVX], vy = v[KVY], vz = v[KVZ], pr = v[KPR], en;
  int err = eos_energy(rh, pr, &en);
  if (err > 0) return err;
  const double *g = G.g_cov;
  double v2 = g[0] * vx * vx + g[1] * vy * vy + g[2] * vz * vz;
  double glf = 1. / sqrt(1. - v2);
  double w = en + pr;
  double ww = w * (glf * glf);
  u[KRH] = rh * glf;
  if (G.viscosity) {
    double pitt, pitx, pity, pitz;
    w = w + v[KPIBU];
    ww = w * (glf * glf);
    derived_pi(v, &pitt, &pitx, &pity, &pitz);
    u[KVX] = (ww * vx + pitx);
    u[KVY] = (ww * vy + pity);
    u[KVZ] = (ww * vz + pitz) * g[2];
    u[KPR] = ww - pr - v[KPIBU] - pitt * G.g_cov0;
    u[KPIBU] = v[KPIBU] * u[KRH];
    u[KPIZZ] = v[KPIZZ] * u[KRH];
    u[KPIXZ] = v[KPIXZ] * u[KRH];
    u[KPIXY] = v[KPIXY] * u[KRH];
    u[KPIYZ] = v[KPIYZ] * u[KRH];
    u[KPIXX] = v[KPIXX] * u[KRH];
    u[KPIYY] = v[KPIYY] * u[KRH];
  } else {
    u[KVX] = ww * vx;
    u[KVY] = ww * vy;
    u[KVZ] = ww * vz * g[2];
    u[KPR] = ww - pr;
  }
  for (int k = 0; k < G.nu; k++) u[k] = u[k] * G.gp;
  return 0;
}

/* system_prim2cons_0, system.f90:267-322 (viscous initialisation) */
int sys_prim2cons_0(double *v, double *u) {
  double rh = v[KRH], vx = v[KVX], vy = v[KVY], vz = v[KVZ], pr = v[KPR], en;
  int err = eos_energy(rh, pr, &en);
  if (err > 0) return err;
  const double *g = G.g_cov;
  double glf = 1. / sqrt(1. - vx * vx * g[0] - vy * vy * g[1] - vz * vz * g[2]);
  double w = en + pr + v[KPIBU];
  double ww = w * (glf * glf);
  u[KRH] = rh * glf;
  double pitt, pitx, pity, pitz;
  derived_pi(v, &pitt, &pitx, &pity, &pitz);
  u[KVX] = ww * vx + pitx;
  u[KVY] = ww * vy + pity;
  u[KVZ] = (ww * vz + pitz) * g[2];
  u[KPR] = ww - pr - v[KPIBU] - pitt * G.g_cov0;
  u[KPIBU] = v[KPIBU] * u[KRH];
  u[KPIZZ] = v[KPIZZ] * u[KRH];
  u[KPIXZ] = v[KPIXZ] * u[KRH];
  u[KPIXY] = v[KPIXY] * u[KRH];
  u[KPIYZ] = v[KPIYZ] * u[KRH];
  u[KPIXX] = v[KPIXX] * u[KRH];
  u[KPIYY] = v[KPIYY] * u[KRH];
  for (int k = 0; k < G.nu; k++) u[k] = u[k] * G.gp;
  return 0;
}

/* system_cons2prim, system.f90:423-665: ideal HD Newton on p (511-562) and
 * viscous nested iteration (564-665) */
int sys_cons2prim(double *u, double *v) {
  const double tol = 1.e-12, eps = 1.e-12;
  const int iter_max = 100, ivmax = 100, ipmax = 100;
  const double *g = G.g_cov;
  const double gm = G.gm;
  int errcode = 0;
  double d = gm * u[KRH];
  double sx = gm * u[KVX];
  double sy = gm * u[KVY];
  double sz = gm * u[KVZ] / g[2];
  double et = gm * u[KPR];
  double s2 = g[0] * sx * sx + g[1] * sy * sy + g[2] * sz * sz;
  double pr, w = 0, v2, glf, rh = 0, en, drhdpr, dendpr, pr_eos = 0, dprdrh, dprden, f, df, dpr;

  if (!G.viscosity && !G.mhd) {
    pr = v[KPR];
    for (int iter = 1; iter <= iter_max; iter++) {
      w = et + pr;
      v2 = s2 / (w * w);
      if (v2 >= 1.) return 777;
      glf = 1. / sqrt(1. - v2);
      rh = d / glf;
      en = w * (1. - v2) - pr;
      drhdpr = rh * v2 * glf / (en + pr);
      dendpr = v2;
      errcode = eos_pressure(rh, en, &pr_eos, &dprdrh, &dprden);
      f = pr_eos - pr;
      df = dprdrh * drhdpr + dprden * dendpr - 1.;
      dpr = -f / df;
      if (fabs(dpr) < tol * pr) break;
      pr = pr + dpr;
    }
    v[KRH] = rh;
    v[KVX] = sx / w;
    v[KVY] = sy / w;
    v[KVZ] = sz / w;
    v[KPR] = pr;
    if (pr <= 0) errcode = 777;
    return errcode;
  }

  if (G.viscosity) {
    double uin[MAXNV], vprim[MAXNV];
    for (int k = 0; k < G.nv; k++) uin[k] = gm * u[k];
    d = uin[KRH];
    for (int k = KPIBU; k <= KPIZZ; k++) v[k] = uin[k] / uin[KRH];
    double vx = v[KVX], vy = v[KVY], vz = v[KVZ];
    v2 = g[0] * (vx * vx) + g[1] * (vy * vy) + g[2] * (vz * vz);
    for (int k = 0; k < G.nv; k++) vprim[k] = v[k];
    for (int iv = 1; iv <= ivmax; iv++) {
      double pitt, pitx, pity, pitz;
      vprim[KVX] = vx; vprim[KVY] = vy; vprim[KVZ] = vz;
      derived_pi(vprim, &pitt, &pitx, &pity, &pitz);
      double en1 = et + v[KPIBU] + pitt * G.g_cov0;
      double sx_cov = uin[KVX] - pitx;
      double sy_cov = uin[KVY] - pity;
      double sz_cov = uin[KVZ] - pitz * g[2];
      sx = sx_cov;
      sy = sy_cov;
      sz = sz_cov / g[2];
      s2 = sx * sx_cov + sy * sy_cov + sz * sz_cov;
      pr = v[KPR];
      for (int ip = 1; ip <= ipmax; ip++) {
        double q = en1 + pr;
        v2 = s2 / (q * q);
        if (v2 >= 1.) return 107;
        glf = 1. / sqrt(1. - v2);
        rh = d / glf;
        en = (en1 + pr) * (1. - v2) - pr - vprim[KPIBU];
        drhdpr = rh * v2 * glf / (en + pr);
        dendpr = v2;
        errcode = eos_pressure(rh, en, &pr_eos, &dprdrh, &dprden);
        f = pr_eos - pr;
        df = dprdrh * drhdpr + dprden * dendpr - 1.;
        dpr = -f / df;
        if (fabs(dpr) < tol * pr) break;
        pr = pr + dpr;
      }
      vx = sx / (en1 + pr); if (fabs(vx) < eps) vx = 0.;
      vy = sy / (en1 + pr); if (fabs(vy) < eps) vy = 0.;
      vz = sz / (en1 + pr); if (fabs(vz) < eps) vz = 0.;
      vprim[KVX] = vx; vprim[KVY] = vy; vprim[KVZ] = vz;
      if (fabs(g[0] * (vx * vx) + g[1] * (vy * vy) + g[2] * (vz * vz) - v2) < eps) break;
    }
    v[KRH] = rh; v[KVX] = vx; v[KVY] = vy; v[KVZ] = vz; v[KPR] = pr;
    if (G.obtained_zz) {
      double pitt, pitx, pity, pitz, pizz;
      get_derived_pi_zz(v, &pitt, &pitx, &pity, &pitz, &pizz);
      v[KPIZZ] = pizz;
    }
    return errcode;
  }
  return 11;   /* MHD not restated in this oracle: exit(11) branch, system.f90:1160 */
}

/* system_flux, system.f90:1316-1583 (HD and viscous branches) */
int sys_flux(double *v, double *u, double *f, double *vf, int dir) {
  double rh = v[KRH], vx = v[KVX], vy = v[KVY], vz = v[KVZ], pr = v[KPR], en;
  int err = eos_energy(rh, pr, &en);
  if (err > 0) return err;
  const double *g = G.g_cov;
  double v2 = g[0] * vx * vx + g[1] * vy * vy + g[2] * vz * vz;
  if (v2 > 1) return 333;
  double glf = 1. / sqrt(1. - v2);
  double w;
  if (!G.viscosity) w = en + pr;
  else w = en + pr + v[KPIBU];
  double ww = w * (glf * glf);
  double pitt = 0, pitx = 0, pity = 0, pitz = 0;
  u[KRH] = rh * glf;
  if (G.viscosity) {
    derived_pi(v, &pitt, &pitx, &pity, &pitz);
    for (int k = KPIBU; k < G.nv; k++) u[k] = v[k] * u[KRH];
    u[KVX] = ww * vx + pitx;
    u[KVY] = ww * vy + pity;
    u[KVZ] = (ww * vz + pitz) * g[2];
    u[KPR] = ww - pr - v[KPIBU] - pitt * G.g_cov0;
  } else {
    u[KVX] = ww * vx;
    u[KVY] = ww * vy;
    u[KVZ] = ww * vz * g[2];
    u[KPR] = ww - pr;
  }
  double wwdir, vdir, gcov_dir;
  if (dir == 1) {
    wwdir = ww * vx; vdir = vx; gcov_dir = g[0];
    f[KRH] = rh * glf * vx;
    if (G.viscosity) {
      f[KVX] = wwdir * vx * g[0] + pr + v[KPIBU] + v[KPIXX] * g[0];
      f[KVY] = wwdir * vy * g[1] + v[KPIXY] * g[1];
      f[KVZ] = wwdir * vz * g[2] + v[KPIXZ] * g[2];
      f[KPR] = wwdir + pitx;
      for (int k = G.nou; k < G.nv; k++) f[k] = u[k] * vx;
    } else {
      f[KVX] = wwdir * vx * g[0] + pr;
      f[KVY] = wwdir * vy * g[1];
      f[KVZ] = wwdir * vz * g[2];
      f[KPR] = wwdir;
    }
  } else if (dir == 2) {
    wwdir = ww * vy; vdir = vy; gcov_dir = g[1];
    f[KRH] = rh * glf * vy;
    if (G.viscosity) {
      f[KVX] = wwdir * vx * g[0] + v[KPIXY] * g[0];
      f[KVY] = wwdir * vy * g[1] + pr + v[KPIBU] + v[KPIYY] * g[1];
      f[KVZ] = wwdir * vz * g[2] + v[KPIYZ] * g[2];
      f[KPR] = wwdir + pity;
      for (int k = G.nou; k < G.nv; k++) f[k] = u[k] * vy;
    } else {
      f[KVX] = wwdir * vx * g[0];
      f[KVY] = wwdir * vy * g[1] + pr;
      f[KVZ] = wwdir * vz * g[2];
      f[KPR] = wwdir;
    }
  } else {
    wwdir = ww * vz; vdir = vz; gcov_dir = g[2];
    f[KRH] = rh * glf * vz;
    if (G.viscosity) {
      f[KVX] = wwdir * vx * g[0] + v[KPIXZ] * g[0];
      f[KVY] = wwdir * vy * g[1] + v[KPIYZ] * g[1];
      f[KVZ] = wwdir * vz * g[2] + pr + v[KPIBU] + v[KPIZZ] * g[2];
      f[KPR] = wwdir + pitz;
      for (int k = G.nou; k < G.nv; k++) f[k] = u[k] * vz;
    } else {
      f[KVX] = wwdir * vx * g[0];
      f[KVY] = wwdir * vy * g[1];
      f[KVZ] = wwdir * vz * g[2] + pr;
      f[KPR] = wwdir;
    }
  }
  for (int k = 0; k < G.nu; k++) { u[k] = u[k] * G.gp; f[k] = f[k] * G.gp; }
  double ca2 = 0., cs2;
  err = eos_sound(rh, en, pr, &cs2);
  if (err > 0) return err;
  double a2 = cs2 + ca2 - cs2 * ca2;
  double den = 1. / (1. - v2 * a2);
  double vf1 = den * (1. - a2) * vdir;
  double vf2 = den * sqrt(a2 * (1. - v2) * ((1. - v2 * a2) / gcov_dir - (1. - a2) * (vdir * vdir)));
  vf[0] = vf1 + vf2;
  vf[1] = vf1 - vf2;
  return 0;
}

/* get_viscous_parameters, viscous.f90:172-199 */
static int get_viscous_parameters(double rh, double pr, double *eta, double *zeta, double *tau_pi, double *tau_pi_big) {
  double en, temp, s, cs2;
  int err = get_derived_data(rh, pr, &en, &temp, &s);
  int err2 = eos_sound(rh, en, pr, &cs2);
  if (err2) err = err2;
  *eta = HBAR * G.eta_over_s * s;
  if (G.init_type == SHEAR_VISCOUS_1D) *eta = G.eta_over_s;
  if (G.bulkvis) *zeta = 2. * *eta * (1. / 3. - cs2);
  else *zeta = 0.;
  if (temp == 0 || s == 0) *tau_pi = 0.;
  else *tau_pi = G.tau_pi_coeff * *eta / (temp * s);
  if (G.init_type == GUBSER_VISCOUS) *tau_pi = 5. * *eta / (temp * s);
  *tau_pi_big = *tau_pi;
  return err;
}

/* get_sigma, viscous.f90:283-393 */
static void get_sigma(int ix, int iy, int iz, double *sigma, double *dd) {
  double ux = V(ix, iy, iz, KVX), uy = V(ix, iy, iz, KVY), uz = V(ix, iy, iz, KVZ);
  const double *g = G.g_cov;
  double ut = 1. / sqrt(1. - g[0] * ux * ux - g[1] * uy * uy - g[2] * uz * uz);
  ux = ux * ut; uy = uy * ut; uz = uz * ut;
  double dutdt = DERIV(ix, iy, iz, DTT), dutdx = DERIV(ix, iy, iz, DTX);
  double dutdy = DERIV(ix, iy, iz, DTY), dutdz = DERIV(ix, iy, iz, DTZ);
  double duxdx = DERIV(ix, iy, iz, DXX), duydy = DERIV(ix, iy, iz, DYY);
  double duzdz = DERIV(ix, iy, iz, DZZ), duzdt = DERIV(ix, iy, iz, DZT);
  double duzdx = DERIV(ix, iy, iz, DZX), duzdy = DERIV(ix, iy, iz, DZY);
  double duxdt = DERIV(ix, iy, iz, DXT), duxdy = DERIV(ix, iy, iz, DXY);
  double duxdz = DERIV(ix, iy, iz, DXZ), duydt = DERIV(ix, iy, iz, DYT);
  double duydx = DERIV(ix, iy, iz, DYX), duydz = DERIV(ix, iy, iz, DYZ);
  /* order of the 16 values handed back to system_source */
  dd[0] = dutdt; dd[1] = dutdx; dd[2] = dutdy; dd[3] = dutdz; dd[4] = duxdx;
  dd[5] = duydy; dd[6] = duzdz; dd[7] = duzdt; dd[8] = duzdx; dd[9] = duzdy;
  dd[10] = duxdt; dd[11] = duxdy; dd[12] = duxdz; dd[13] = duydt; dd[14] = duydx; dd[15] = duydz;
  double ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
  if (G.coordinates == MINKOWSKI) {
    sigma[SZZ] = ut * uz * duzdt + uz * ux * duzdx + uz * uy * duzdy + (2. / 3.) * duzdz + (2. / 3.) * duzdz * uz2 - dutdt / 3. - duxdx / 3. - duydy / 3. - uz2 * dutdt / 3.
                 - uz2 * duxdx / 3. - uz2 * duydy / 3.;
    sigma[SXZ] = 0.5 * ut * ux * duzdt + 0.5 * duzdx + 0.5 * ux2 * duzdx + 0.5 * ux * uy * duzdy + ux * uz * duzdz / 6. + 0.5 * ut * uz * duxdt + ux * uz * duxdx / 6. + 0.5 * uy * uz *
                 duxdy + 0.5 * duxdz + 0.5 * duxdz * uz2 - ux * uz * dutdt / 3. - ux * uz * duydy / 3.;
    sigma[SXY] = 0.5 * ut * ux * duydt + 0.5 * duydx + 0.5 * ux2 * duydx + ux * uy * duydy / 6. + ux * uz * duydz / 2. + 0.5 * ut * uy * duxdt + ux * uy * duxdx / 6. + 0.5 * duxdy +
                 0.5 * duxdy * uy2 + uy * uz * duxdz / 2. - ux * uy * dutdt / 3. - ux * uy * duzdz / 3.;
    sigma[SYZ] = 0.5 * ut * uy * duzdt + ux * uy * duzdx / 2. + duzdy / 2. + duzdy * uy2 / 2. + uy * uz * duzdz / 6. + ut * uz * duydt / 2. + ux * uz * duydx / 2. + uy * uz * duydy / 6.
                 + duydz / 2. + duydz / 2. * uz2 - uy * uz * dutdt / 3. - uy * uz * duxdx / 3.;
    sigma[SXX] = ut * ux * duxdt + (2. / 3.) * duxdx + (2. / 3.) * duxdx * ux2 + ux * uy * duxdy + ux * uz * duxdz - dutdt / 3. - duydy / 3. - duzdz / 3. - ux2 *
                 dutdt / 3. - ux2 * duydy / 3. - ux2 * duzdz / 3.;
    sigma[SYY] = ut * uy * duydt + ux * uy * duydx + (2. / 3.) * duydy + (2. / 3.) * duydy * uy2 + uy * uz * duydz - dutdt / 3. - duxdx / 3. - duzdz / 3. - uy2 *
                 dutdt / 3. - uy2 * duxdx / 3. - uy2 * duzdz / 3.;
  } else {
    const double t = G.t;
    const double t2 = pow(t, 2.), t3 = pow(t, 3.);
    sigma[SZZ] = ut * uz * duzdt + uz * ux * duzdx + uz * uy * duzdy + (2. * duzdz) / (3. * t2) + (2. / 3.) * duzdz * uz2 - dutdt / (3. * t2) - duxdx / (3. *
                 t2) - duydy / (3. * t2) - uz2 * dutdt / 3. - uz2 * duxdx / 3. - uz2 * duydy / 3.
                 + 2. / 3. * ut / t3 + 5. / 3. * (ut / t) * uz2;
    sigma[SXZ] = 0.5 * ut * ux * duzdt + 0.5 * duzdx + 0.5 * ux2 * duzdx + 0.5 * ux * uy * duzdy + ux * uz * duzdz / 6. + 0.5 * ut * uz * duxdt + ux * uz * duxdx / 6. + 0.5 * uy * uz *
                 duxdy + duxdz / (2. * t2) + 0.5 * duxdz * uz2 - ux * uz * dutdt / 3. - ux * uz * duydy / 3. + (2. / 3.) * ut * uz * ux / t;
    sigma[SXY] = 0.5 * ut * ux * duydt + 0.5 * duydx + 0.5 * ux2 * duydx + ux * uy * duydy / 6. + ux * uz * duydz / 2. + 0.5 * ut * uy * duxdt + ux * uy * duxdx / 6. + 0.5 * duxdy +
                 0.5 * duxdy * uy2 + uy * uz * duxdz / 2. - ux * uy * dutdt / 3. - ux * uy * duzdz / 3. - ut * ux * uy / (3. * t);
    sigma[SYZ] = 0.5 * ut * uy * duzdt + ux * uy * duzdx / 2. + duzdy / 2. + duzdy * uy2 / 2. + uy * uz * duzdz / 6. + ut * uz * duydt / 2. + ux * uz * duydx / 2. + uy * uz * duydy / 6.
                 + duydz / (2. * t2) + duydz / 2. * uz2 - uy * uz * dutdt / 3. - uy * uz * duxdx / 3. + (2. / 3.) * ut * uz * uy / t;
    sigma[SXX] = ut * ux * duxdt + (2. / 3.) * duxdx + (2. / 3.) * duxdx * ux2 + ux * uy * duxdy + ux * uz * duxdz - dutdt / 3. - duydy / 3. - duzdz / 3. - ut / (3. * t) -
                 ux2 * dutdt / 3. - ux2 * duydy / 3. - ux2 * duzdz / 3. - ux2 * ut / (3. * t);
    sigma[SYY] = ut * uy * duydt + ux * uy * duydx + (2. / 3.) * duydy + (2. / 3.) * duydy * uy2 + uy * uz * duydz - dutdt / 3. - duxdx / 3. - duzdz / 3. - ut / (3. * t) -
                 uy2 * dutdt / 3. - uy2 * duxdx / 3. - uy2 * duzdz / 3. - uy2 * ut / (3. * t);
  }
}

/* system_source, system.f90:1587-1842 (HD and viscous branches) */
int sys_source(int ix, int iy, int iz, double *v, double *src, double *src_stiff) {
  const double lambda0 = 1., lambda1 = 1., lambda2 = 0.;
  for (int k = 0; k < G.nv; k++) { src[k] = 0.; src_stiff[k] = 0.; }
  double rh = v[KRH], vx = v[KVX], vy = v[KVY], vz = v[KVZ], pr = v[KPR], en;
  int err = eos_energy(rh, pr, &en);
  if (err > 0) return err;
  const double *g = G.g_cov;
  const double t = G.t;
  double v2 = g[0] * vx * vx + g[1] * vy * vy + g[2] * vz * vz;
  double glf = 1. / sqrt(1. - v2);
  double ut = glf;
  if (!G.viscosity) {
    if (G.coordinates == BJORKEN) {
      double w = en + pr;
      double ww = w * (glf * glf);
      src[KPR] = -t * (ww * vz * vz + pr / g[2]);
    }
  } else {
    double sigma[6], dd[16], eta, zeta, tau_pi, tau_pi_big;
    get_sigma(ix, iy, iz, sigma, dd);
    err = get_viscous_parameters(v[KRH], v[KPR], &eta, &zeta, &tau_pi, &tau_pi_big);
    if (err > 0) return err;
    double pitt, pitx, pity, pitz;
    derived_pi(v, &pitt, &pitx, &pity, &pitz);
    double dutdt = dd[0], dutdx = dd[1], dutdy = dd[2], dutdz = dd[3], duxdx = dd[4];
    double duydy = dd[5], duzdz = dd[6], duzdt = dd[7], duzdx = dd[8], duzdy = dd[9];
    double duxdt = dd[10], duxdy = dd[11], duxdz = dd[12], duydt = dd[13], duydx = dd[14], duydz = dd[15];
    double sum_der = dutdt + duxdx + duydy + duzdz;
    double sum_der_0, sum_der_x, sum_der_y, sum_der_z, sdt;
    double *p = v;
    if (G.coordinates == BJORKEN) {
      double sum_der_t = sum_der + ut / t;
      sdt = sum_der_t;
      sum_der_0 = G.g_cov0 * (glf * dutdt + glf * p[KVX] * dutdx + glf * p[KVY] * dutdy + glf * p[KVZ] * dutdz);
      sum_der_x = g[0] * (glf * duxdt + glf * p[KVX] * duxdx + glf * p[KVY] * duxdy + glf * p[KVZ] * duxdz);
      sum_der_y = g[1] * (glf * duydt + glf * p[KVX] * duydx + glf * p[KVY] * duydy + glf * p[KVZ] * duydz);
      sum_der_z = glf * (pow(t, 2.) * duzdt + 2. * t * p[KVZ]) + pow(t, 2.) * (glf * p[KVX] * duzdx + glf * p[KVY] * duzdy + glf * p[KVZ] * duzdz);
      double w = en + pr + p[KPIBU];
      double ww = w * (glf * glf);
      src[KPR] = -t * (ww * vz * vz + (pr + p[KPIBU]) / g[2] + p[KPIZZ]);
      src[KPIBU] = -1. / tau_pi_big * (p[KPIBU] + zeta * sdt) - lambda0 * 4. / 3. * p[KPIBU] * sdt;
    } else {
      sdt = sum_der;
      sum_der_0 = G.g_cov0 * (glf * dutdt + p[KVX] * dutdx + p[KVY] * dutdy + p[KVZ] * dutdz);
      sum_der_x = g[0] * (glf * duxdt + p[KVX] * duxdx + p[KVY] * duxdy + p[KVZ] * duxdz);
      sum_der_y = g[1] * (glf * duydt + p[KVX] * duydx + p[KVY] * duydy + p[KVZ] * duydz);
      sum_der_z = g[2] * (glf * duzdt + p[KVX] * duzdx + p[KVY] * duzdy + p[KVZ] * duzdz);
      src[KPIBU] = -1. / tau_pi_big * (p[KPIBU] + zeta * sum_der) + lambda0 * (-1. / 3. * p[KPIBU] * sum_der);
    }
    double w0x = -0.5 * (dutdx + duxdt) + 0.5 * glf * (sum_der_x - p[KVX] * g[0] * sum_der_0);
    double w0y = -0.5 * (dutdy + duydt) + 0.5 * glf * (sum_der_y - p[KVY] * g[1] * sum_der_0);
    double w0z = -0.5 * (dutdz + duzdt) + 0.5 * glf * (sum_der_z - p[KVZ] * g[2] * sum_der_0);
    double wxy = 0.5 * (g[1] * duydx - duxdy) + 0.5 * glf * (g[0] * p[KVX] * sum_der_y - g[1] * p[KVY] * sum_der_x);
    double wxz = 0.5 * (g[2] * duzdx - duxdz) + 0.5 * glf * (g[0] * p[KVX] * sum_der_z - g[2] * p[KVZ] * sum_der_x);
    double wyz = 0.5 * (g[2] * duzdy - duydz) + 0.5 * glf * (g[1] * p[KVY] * sum_der_z - g[2] * p[KVZ] * sum_der_y);
    double wx0 = w0x, wy0 = w0y, wz0 = w0z, wyx = -wxy, wzx = -wxz, wzy = -wyz;
    double IVxx = 2. * (pitx * wx0 + p[KPIXY] * wxy + p[KPIXZ] * wxz);
    double IVyy = 2. * (pity * wy0 + p[KPIXY] * wyx + p[KPIYZ] * wyz);
    double IVzz = 2. * (pitz * wz0 + p[KPIXZ] * wzx + p[KPIYZ] * wzy);
    double IVxy = p[KPIXX] * wyx + p[KPIYY] * wxy + p[KPIXZ] * wyz + p[KPIYZ] * wxz + pitx * wy0 + pity * wx0;
    double IVxz = p[KPIXX] * wzx + p[KPIZZ] * wxz + p[KPIXY] * wzy + p[KPIYZ] * wxy + pitx * wz0 + pitz * wx0;
    double IVyz = p[KPIYY] * wzy + p[KPIZZ] * wyz + p[KPIXY] * wzx + p[KPIXZ] * wyx + pity * wz0 + pitz * wy0;
    if (G.coordinates == BJORKEN) {
      double sum_der_t = sdt;
      src[KPIZZ] = -1. / tau_pi * (p[KPIZZ] + 2. * eta * sigma[SZZ]) - 2. * glf * p[KPIZZ] / t -
                   2. * (glf * vz) * pitz / t - lambda0 * 4. / 3. * p[KPIZZ] * sum_der_t + lambda1 * glf * (2. * p[KVZ] * (pitz * sum_der_0 +
                   p[KPIXZ] * sum_der_x + p[KPIYZ] * sum_der_y + p[KPIZZ] * sum_der_z)) - lambda2 * IVzz;
      src[KPIXZ] = -1. / tau_pi * (p[KPIXZ] + 2. * eta * sigma[SXZ]) + lambda0 * (-4. / 3. * p[KPIXZ] * sum_der_t) - glf / t * p[KPIXZ] - glf * vz * pitx / t +
                   lambda1 * glf * ((pitx * p[KVZ] + pitz * p[KVX]) * sum_der_0 + (p[KPIXX] * p[KVZ] + p[KPIXZ] * p[KVX]) * sum_der_x + (p[KPIXY] * p[KVZ] +
                   p[KPIYZ] * p[KVX]) * sum_der_y + (p[KPIXZ] * p[KVZ] + p[KPIZZ] * p[KVX]) * sum_der_z) - lambda2 * IVxz;
      src[KPIYZ] = -1. / tau_pi * (p[KPIYZ] + 2. * eta * sigma[SYZ]) + lambda0 * (-4. / 3. * p[KPIYZ] * sum_der_t) - glf / t * p[KPIYZ] - glf * vz * pity / t +
                   lambda1 * glf * ((pity * p[KVZ] + pitz * p[KVY]) * sum_der_0 + (p[KPIXY] * p[KVZ] + p[KPIXZ] * p[KVY]) * sum_der_x + (p[KPIYY] * p[KVZ] +
                   p[KPIYZ] * p[KVY]) * sum_der_y + (p[KPIYZ] * p[KVZ] + p[KPIZZ] * p[KVY]) * sum_der_z) - lambda2 * IVyz;
      src[KPIXY] = -1. / tau_pi * (p[KPIXY] + 2. * eta * sigma[SXY]) + lambda0 * (-4. / 3. * p[KPIXY] * sum_der_t) + lambda1 * glf * ((pitx *
                   p[KVY] + pity * p[KVX]) * sum_der_0 + (p[KPIXX] * p[KVY] + p[KPIXY] * p[KVX]) * sum_der_x + (p[KPIXY] * p[KVY] + p[KPIYY] * p[KVX]) *
                   sum_der_y + (p[KPIXZ] * p[KVY] + p[KPIYZ] * p[KVX]) * sum_der_z) - lambda2 * IVxy;
      src[KPIXX] = -1. / tau_pi * (p[KPIXX] + 2. * eta * sigma[SXX]) + lambda0 * (-4. / 3. * p[KPIXX] * sum_der_t) + lambda1 * glf * (2. * p[KVX] * (pitx *
                   sum_der_0 + p[KPIXX] * sum_der_x + p[KPIXY] * sum_der_y + p[KPIXZ] * sum_der_z)) - lambda2 * IVxx;
      src[KPIYY] = -1. / tau_pi * (p[KPIYY] + 2. * eta * sigma[SYY]) + lambda0 * (-4. / 3. * p[KPIYY] * sum_der_t) + lambda1 * glf * (2. * p[KVY] * (pity *
                   sum_der_0 + p[KPIXY] * sum_der_x + p[KPIYY] * sum_der_y + p[KPIYZ] * sum_der_z)) - lambda2 * IVyy;
    } else {
      src[KPIZZ] = -1. / tau_pi * (p[KPIZZ] + 2. * eta * sigma[SZZ]) + lambda0 * (-4. / 3. * p[KPIZZ] * sum_der) + lambda1 * (2. * p[KVZ] * (pitz * sum_der_0
                   + p[KPIXZ] * sum_der_x + p[KPIYZ] * sum_der_y + p[KPIZZ] * sum_der_z)) - lambda2 * IVzz;
      src[KPIXZ] = -1. / tau_pi * (p[KPIXZ] + 2. * eta * sigma[SXZ]) + lambda0 * (-4. / 3. * p[KPIXZ] * sum_der) + lambda1 * ((pitx * p[KVZ] + pitz * p[KVX]) *
                   sum_der_0 + (p[KPIXX] * p[KVZ] + p[KPIXZ] * p[KVX]) * sum_der_x + (p[KPIXY] * p[KVZ] + p[KPIYZ] * p[KVX]) * sum_der_y + (p[KPIXZ] * p[KVZ]
                   + p[KPIZZ] * p[KVX]) * sum_der_z) - lambda2 * IVxz;
      src[KPIXY] = -1. / tau_pi * (p[KPIXY] + 2. * eta * sigma[SXY]) + lambda0 * (-4. / 3. * p[KPIXY] * sum_der) + lambda1 * ((pitx * p[KVY] + pity * p[KVX]) *
                   sum_der_0 + (p[KPIXX] * p[KVY] + p[KPIXY] * p[KVX]) * sum_der_x + (p[KPIXY] * p[KVY] + p[KPIYY] * p[KVX]) * sum_der_y + (p[KPIXZ] * p[KVY]
                   + p[KPIYZ] * p[KVX]) * sum_der_z) - lambda2 * IVxy;
      src[KPIYZ] = -1. / tau_pi * (p[KPIYZ] + 2. * eta * sigma[SYZ]) + lambda0 * (-4. / 3. * p[KPIYZ] * sum_der) + lambda1 * ((pity * p[KVZ] + pitz * p[KVY]) *
                   sum_der_0 + (p[KPIXY] * p[KVZ] + p[KPIXZ] * p[KVY]) * sum_der_x + (p[KPIYY] * p[KVZ] + p[KPIYZ] * p[KVY]) * sum_der_y + (p[KPIYZ] * p[KVZ]
                   + p[KPIZZ] * p[KVY]) * sum_der_z) - lambda2 * IVyz;
      src[KPIXX] = -1. / tau_pi * (p[KPIXX] + 2. * eta * sigma[SXX]) + lambda0 * (-4. / 3. * p[KPIXX] * sum_der) + lambda1 * (2. * p[KVX] * (pitx * sum_der_0 +
                   p[KPIXX] * sum_der_x + p[KPIXY] * sum_der_y + p[KPIXZ] * sum_der_z)) - lambda2 * IVxx;
      src[KPIYY] = -1. / tau_pi * (p[KPIYY] + 2. * eta * sigma[SYY]) + lambda0 * (-4. / 3. * p[KPIYY] * sum_der) + lambda1 * (2. * p[KVY] * (pity * sum_der_0 +
                   p[KPIXY] * sum_der_x + p[KPIYY] * sum_der_y + p[KPIYZ] * sum_der_z)) - lambda2 * IVyy;
    }
    for (int k = KPIBU; k < G.nv; k++) src[k] = src[k] * v[KRH];
  }
  for (int k = 0; k < G.nv; k++) src[k] = src[k] * G.gp;
  return 0;
}
