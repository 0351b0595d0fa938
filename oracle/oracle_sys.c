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

/* system_compute_E_in_ideal_MHD, system.f90:1882-1897: covariant E = -v x B
 * (times sqrt(-g)) from contravariant v and B */
static void compute_E_ideal(const double *V, const double *B, double *E) {
  E[0] = -G.gp * (V[1] * B[2] - V[2] * B[1]);
  E[1] = -G.gp * (V[2] * B[0] - V[0] * B[2]);
  E[2] = -G.gp * (V[0] * B[1] - V[1] * B[0]);
}

/* system_EXB, system.f90:1846-1862: E x B from covariant E, contravariant B */
static void system_EXB(const double *E, const double *B, double *X) {
  double c0 = E[0], c1 = E[1], c2 = E[2] / G.g_cov[2];
  X[0] = G.gp * (c1 * B[2] - c2 * B[1]);
  X[1] = G.gp * (c2 * B[0] - c0 * B[2]);
  X[2] = G.gp * (c0 * B[1] - c1 * B[0]);
}

/* system_prim2cons, system.f90:164-263 (HD, viscous and ideal-MHD branches) */
int sys_prim2cons(double *v, double *u) {
  if (G.rmhd) return rmhd_prim2cons(v, u, 0);
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
  } else if (G.mhd) {
    double E[3], EXB[3];
    compute_E_ideal(&v[KVX], &v[G.kbx], E);
    double E2 = E[0] * E[0] + E[1] * E[1] + E[2] * E[2] / g[2];
    system_EXB(E, &v[G.kbx], EXB);
    double B2 = v[G.kbx] * v[G.kbx] + v[G.kby] * v[G.kby] + v[G.kbz] * v[G.kbz] * g[2];
    double EEM = 0.5 * (E2 + B2);
    u[KVX] = ww * vx + EXB[0];
    u[KVY] = ww * vy + EXB[1];
    u[KVZ] = ww * vz * g[2] + EXB[2];
    u[KPR] = ww - pr + EEM;
    u[G.kbx] = v[G.kbx]; u[G.kby] = v[G.kby]; u[G.kbz] = v[G.kbz];
    if (G.divclean) u[G.kglm] = v[G.kglm] / G.gp;
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

/* ---------------------------------------------------------------------- */
/* Ideal-MHD con2prim: algorithms 4, 5 and 6 of system_cons2prim
 * (system.f90:880-1158) with rtsafe (2242-2290) and the cubic solvers
 * solver7 / solver8 (2102-2201).  solver_par = (B2, SB2, S2, U, D) as the
 * reference's solver_parameters; w_solver is the W of the last cubic solve
 * (the reference's w_algo_solver7 / w_algo_solver8_e_3p). */
static double solver_par[5];
static double w_solver;

/* solver7, system.f90:2102-2148 (ideal gas, Gamma = 4/3) */
static void solver7(double x, double *f, double *df) {
  const double third = 1.0 / 3.0;
  const double gamma_ad = 4.0 / 3.0;
  double gam = (gamma_ad - 1) / gamma_ad;
  double B2 = solver_par[0], SB2 = solver_par[1], S2 = solver_par[2], UU = solver_par[3], D = solver_par[4];
  double v2 = x;
  double rho = D * sqrt(1.0 - v2);
  double c3 = 1.0 - gam * (1.0 - v2);
  double c2 = gam * rho + .5 * B2 * (1.0 + v2) - UU;
  double c0 = -0.5 * SB2;
  double w = fmax2(-c2 / c3, pow(-c0 / c3, third));
  for (int iter = 1; iter <= 50; iter++) {
    double dw = -((c3 * w + c2) * (w * w) + c0) / ((3 * c3 * w + 2 * c2) * w);
    if (fabs(dw / w) < 1.e-10) break;
    w = w + dw;
  }
  double dc3 = gam;
  double dc2 = .5 * (B2 - gam * rho / (1. - v2));
  double dlogw = -(dc3 * w + dc2) / (3 * c3 * w + 2 * c2);
  double wb = w + B2;
  double vb2 = SB2 / (w * w);
  *f = (wb * wb) * v2 - (2 * w + B2) * vb2 - S2;
  *df = wb * (wb + 2 * dlogw * (w * v2 + vb2));
  w_solver = w;
}

/* solver8, system.f90:2152-2201 (p = e/3) */
static void solver8(double x, double *f, double *df) {
  const double third = 1.0 / 3.0;
  const double gam = 0.25;
  double B2 = solver_par[0], SB2 = solver_par[1], S2 = solver_par[2], UU = solver_par[3];
  double v2 = x;
  double c3 = 1.0 - gam * (1.0 - v2);
  double c2 = .5 * B2 * (1.0 + v2) - UU;
  double c0 = -0.5 * SB2;
  double w;
  if (c0 == 0.0) {
    w = -c2 / c3;
  } else {
    w = fmax2(-c2 / c3, pow(-c0 / c3, third));
    for (int iter = 1; iter <= 100; iter++) {
      double dw = -((c3 * w + c2) * (w * w) + c0) / ((3 * c3 * w + 2 * c2) * w);
      if (fabs(dw / w) < 1.e-12) break;
      w = w + dw;
    }
  }
  double dc3 = gam;
  double dc2 = .5 * B2;
  double dlogw = -(dc3 * w + dc2) / (3 * c3 * w + 2 * c2);
  double wb = w + B2;
  double vb2 = SB2 / (w * w);
  *f = (wb * wb) * v2 - (2 * w + B2) * vb2 - S2;
  *df = wb * (wb + 2 * dlogw * (w * v2 + vb2));
  w_solver = w;
}

/* rtsafe, system.f90:2242-2290 (Newton safeguarded by bisection) */
static double rtsafe(void (*func)(double, double *, double *), double x0, double x1, double x2,
                     double xacc, int *ierr) {
  const int maxit = 100;
  double f, df, xl = x1, xh = x2, r = x0, dxold = fabs(x2 - x1), dx = dxold, temp;
  *ierr = 0;
  func(r, &f, &df);
  for (int j = 1; j <= maxit; j++) {
    if (((r - xh) * df - f) * ((r - xl) * df - f) > 0. || fabs(2. * f) > fabs(dxold * df)) {
      dxold = dx;
      dx = 0.5 * (xh - xl);
      r = xl + dx;
      if (xl == r) { func(r, &f, &df); return r; }
    } else {
      dxold = dx;
      dx = f / df;
      temp = r;
      r = r - dx;
      if (temp == r) { func(r, &f, &df); return r; }
    }
    if (fabs(dx) < xacc) return r;
    func(r, &f, &df);
    if (f < 0.) xl = r; else xh = r;
  }
  *ierr = 1;
  return r;
}

/* get_x9, system.f90:2294-2301 */
static double get_x9(double y, double s2, double b2, double sb2) {
  return (s2 * (y * y) + b2 * sb2 + 2 * sb2 * y) / ((y * y) * ((b2 * b2) + 2 * b2 * y + (y * y)));
}

/* solver10_approximate, system.f90:2205-2223 */
static int solver10_approximate(double b, double c, double d, double oldsol, double *x) {
  double xx = oldsol;
  int iter;
  for (iter = 1; iter <= 100; iter++) {
    double dx = -(12 * (xx * xx * xx) + 3 * b * (xx * xx) + 2 * c * xx + d);
    if (fabs(dx / xx) < 1.e-12) break;
    xx = xx + dx;
  }
  *x = xx;
  return iter == 100 ? 12 : 0;
}

/* ---------------------------------------------------------------------- */
/* BRENTM: Brent's method for n nonlinear equations (ACM algorithm 554,
 * More and Cosnard), restated from libacm554.f90:8-359 as the reference
 * vendors it.  fcn(x, f, k) evaluates component k (0-based) only.  q is
 * column-major with leading dimension n.  Returns INFO. */
typedef void (*brent_fcn)(const double *x, double *f, int k);

static int brentm(brent_fcn fcn, int n, double *x, double *fvec, double ftol, double xtol, int maxfev,
                  int mopt, double *q, double *sigma, double *wa1, double *wa2) {
  const double p05 = 5.e-2, scale = 1.e1;
  const double epsmch = 2.220446049250313e-16;   /* EPSILON(1.d0) */
  int info = 0, nfcall = 0, nfev = 0;
  int nier6 = -1, nier7 = -1, nier8 = 0;
  double fnorm = 0., difit = 0., xnorm = 0.;
  for (int i = 0; i < n; i++) xnorm = fmax2(xnorm, fabs(x[i]));
  const double eps = sqrt(epsmch);
  double delta = scale * xnorm;
  if (xnorm == 0.) delta = scale;
  for (;;) {                                   /* label 20 */
    int nsing = n;
    double fnorm1 = fnorm, difit1 = difit;
    fnorm = 0.;
    double h = eps * xnorm;
    if (h == 0.) h = eps;
    for (int j = 0; j < n; j++) {
      for (int i = 0; i < n; i++) q[i + j * n] = 0.;
      q[j + j * n] = h;
      wa1[j] = x[j];
    }
    for (int k = 0; k < n; k++) {              /* subiteration, label 150 */
      fcn(wa1, fvec, k);
      const double fky = fvec[k];
      nfcall++;
      nfev = nfcall / n;
      fnorm = fmax2(fnorm, fabs(fky));
      for (int j = k; j < n; j++) {
        for (int i = 0; i < n; i++) wa2[i] = wa1[i] + q[i + j * n];
        fcn(wa2, fvec, k);
        const double fkz = fvec[k];
        nfcall++;
        nfev = nfcall / n;
        sigma[j] = fkz - fky;
      }
      fvec[k] = fky;
      double eta = 0.;
      for (int i = k; i < n; i++) eta = fmax2(eta, fabs(sigma[i]));
      if (eta == 0.) continue;
      nsing = nsing - 1;
      double sknorm = 0.;
      for (int i = k; i < n; i++) {
        sigma[i] = sigma[i] / eta;
        sknorm = sknorm + sigma[i] * sigma[i];
      }
      sknorm = sqrt(sknorm);
      if (sigma[k] < 0.) sknorm = -sknorm;
      sigma[k] = sigma[k] + sknorm;
      for (int i = 0; i < n; i++) wa2[i] = 0.;
      for (int j = k; j < n; j++) {
        const double temp = sigma[j];
        for (int i = 0; i < n; i++) wa2[i] = wa2[i] + temp * q[i + j * n];
      }
      for (int j = k; j < n; j++) {
        const double temp = sigma[j] / (sknorm * sigma[k]);
        for (int i = 0; i < n; i++) q[i + j * n] = q[i + j * n] - temp * wa2[i];
      }
      sigma[k] = sknorm * eta;
      double temp = fky / sigma[k];
      if (h * fabs(temp) > delta) temp = copysign(delta / h, temp);
      for (int i = 0; i < n; i++) wa1[i] = wa1[i] + temp * q[i + k * n];
    }
    xnorm = 0.;
    difit = 0.;
    for (int i = 0; i < n; i++) {
      xnorm = fmax2(xnorm, fabs(wa1[i]));
      difit = fmax2(difit, fabs(x[i] - wa1[i]));
      x[i] = wa1[i];
    }
    delta = fmax2(delta, scale * xnorm);
    const int conv = (fnorm < fnorm1 && difit < difit1 && nsing == 0);
    nier6++;
    nier7++;
    nier8++;
    if (conv) nier6 = 0;
    if (fnorm < fnorm1 || difit < difit1) nier7 = 0;
    if (difit > eps * xnorm) nier8 = 0;
    if (fnorm <= ftol) info = 1;
    if (difit <= xtol * xnorm && conv) info = 2;
    if (fnorm <= ftol && info == 2) info = 3;
    if (info != 0) return info;
    if (nfev >= maxfev) info = 4;
    if (nsing == n) info = 5;
    if (nier6 == 5) info = 6;
    if (nier7 == 3) info = 7;
    if (nier8 == 4) info = 8;
    if (info != 0) return info;
    if (!conv || difit > p05 * xnorm || mopt == 1) continue;
    for (int m = 2; m <= mopt; m++) {          /* iterative refinement */
      fnorm1 = fnorm;
      fnorm = 0.;
      int stop = 0;
      for (int k = 0; k < n; k++) {
        fcn(wa1, fvec, k);
        const double fky = fvec[k];
        nfcall++;
        nfev = nfcall / n;
        fnorm = fmax2(fnorm, fabs(fky));
        if (!(fnorm < fnorm1)) {
          fnorm = fnorm1;
          stop = 1;
          break;
        }
        const double temp = fky / sigma[k];
        for (int i = 0; i < n; i++) wa1[i] = wa1[i] + temp * q[i + k * n];
      }
      if (stop) break;
      xnorm = 0.;
      difit = 0.;
      for (int i = 0; i < n; i++) {
        xnorm = fmax2(xnorm, fabs(wa1[i]));
        difit = fmax2(difit, fabs(x[i] - wa1[i]));
        x[i] = wa1[i];
      }
      if (fnorm <= ftol) info = 1;
      if (difit <= xtol * xnorm) info = 2;
      if (fnorm <= ftol && info == 2) info = 3;
      if (nfev >= maxfev && info == 0) info = 4;
      if (info != 0) return info;
    }
  }
}

/* sist3 / sist4, system.f90:1199-1266: the 3x3 (v^2, W, p) and 2x2 (W, p)
 * systems of algorithms 1 and 2.  xinner is the v^2 of the last sist4 call
 * (a module variable in the reference, read after BRENTM returns).  An EoS
 * error inside leaves en as it was (the reference leaves it undefined). */
static double xinner;

static void sist3(const double *X, double *F, int k) {
  const double xx = X[0], yy = X[1], pp = X[2];
  const double B2 = solver_par[0], SB2 = solver_par[1], S2 = solver_par[2], UU = solver_par[3], D = solver_par[4];
  if (k == 0) F[0] = (yy + B2) * (yy + B2) * xx - SB2 / (yy * yy) * (2 * yy + B2) - S2;
  if (k == 1) F[1] = yy - pp + 0.5 * (1. + xx) * B2 - 0.5 * SB2 / (yy * yy) - UU;
  if (k == 2) {
    double rh = D * sqrt(1 - xx), en = 0.;
    eos_energy(rh, pp, &en);
    F[2] = yy - (pp + en) / (1 - xx);
  }
}

static void sist4(const double *X, double *F, int k) {
  const double yy = X[0], pp = X[1];
  const double B2 = solver_par[0], SB2 = solver_par[1], S2 = solver_par[2], UU = solver_par[3], D = solver_par[4];
  const double xx = (S2 + SB2 * (2. * yy + B2) / (yy * yy)) / ((yy + B2) * (yy + B2));
  xinner = xx;
  if (k == 1) {
    double rh = D * sqrt(1 - xx), en = 0.;
    eos_energy(rh, pp, &en);
    F[1] = yy - (pp + en) / (1 - xx);
  }
  if (k == 0) F[0] = yy - pp + 0.5 * (1. + xx) * B2 - 0.5 * SB2 / (yy * yy) - UU;
}

static int mhd_cons2prim(double *u, double *v) {
  const double eps = 1.e-12;
  const double *g = G.g_cov;
  const double gm = G.gm;
  const int kbx = G.kbx, kby = G.kby, kbz = G.kbz;
  double uin[MAXNV];
  int ierr = 0;
  if (G.algo_solver == 1 || G.algo_solver == 2) {
    /* system.f90:668-793: BRENTM on sist3 (3x3) or sist4 (2x2), warm start
     * from the cell's previous primitives */
    for (int k = 0; k < G.nv; k++) uin[k] = gm * u[k];
    double DD = uin[KRH], sx = uin[KVX], sy = uin[KVY], sz = uin[KVZ], UU = uin[KPR];
    double bx = uin[kbx], by = uin[kby], bz = uin[kbz];
    v[kbx] = uin[kbx]; v[kby] = uin[kby]; v[kbz] = uin[kbz];
    double B2 = bx * bx + by * by + bz * bz * g[2];
    double SB = sx * bx + sy * by + sz * bz;
    double SB2 = SB * SB;
    double SS2 = sx * sx + sy * sy + sz * sz / g[2];
    double rh = v[KRH], vx = v[KVX], vy = v[KVY], vz = v[KVZ], pr = v[KPR], en = 0.;
    /* the error code of this lookup is the routine's error code (the
     * reference passes its errcode argument to eos_energy here) */
    const int err0 = eos_energy(rh, pr, &en);
    double v2 = g[0] * (vx * vx) + g[1] * (vy * vy) + g[2] * (vz * vz);
    solver_par[0] = B2; solver_par[1] = SB2; solver_par[2] = SS2; solver_par[3] = UU; solver_par[4] = DD;
    double q[9], sigma[3], wa1[3], wa2[3], fv[3], X[3];
    double xs, ys;
    if (G.algo_solver == 1) {
      X[0] = v2;
      X[1] = (en + pr) / (1 - v2);
      X[2] = pr;
      brentm(sist3, 3, X, fv, 1.e-14, 1.e-14, 300, 3, q, sigma, wa1, wa2);
      xs = X[0]; ys = X[1]; pr = X[2];
    } else {
      X[0] = (en + pr) / (1 - v2);
      X[1] = pr;
      brentm(sist4, 2, X, fv, 1.e-14, 1.e-14, 300, 3, q, sigma, wa1, wa2);
      ys = X[0]; pr = X[1]; xs = xinner;
    }
    rh = DD * sqrt(1 - xs);
    vx = (sx + SB * bx / ys) / (ys + B2);
    vy = (sy + SB * by / ys) / (ys + B2);
    vz = (sz / g[2] + SB * bz / ys) / (ys + B2);
    v[KRH] = rh; v[KVX] = vx; v[KVY] = vy; v[KVZ] = vz; v[KPR] = pr;
    if (G.divclean) v[G.kglm] = u[G.kglm];
    return err0;
  }
  if (G.algo_solver == 4 || G.algo_solver == 5) {
    /* system.f90:880-998 */
    for (int k = 0; k < G.nv; k++) uin[k] = gm * u[k];
    double D = uin[KRH], sx = uin[KVX], sy = uin[KVY], sz = uin[KVZ], UU = uin[KPR];
    double bx = uin[kbx], by = uin[kby], bz = uin[kbz];
    double S2 = sx * sx + sy * sy + sz * sz / g[2];
    double B2 = bx * bx + by * by + bz * bz * g[2];
    double SB = (sx * bx + sy * by + sz * bz);
    double SB2 = SB * SB;
    solver_par[0] = B2; solver_par[1] = SB2; solver_par[2] = S2; solver_par[3] = UU; solver_par[4] = D;
    double v2, w;
    if (G.algo_solver == 4) {
      double gam = (4.0 / 3.0 - 1) / (4.0 / 3.0);
      double v2_old = v[KVX] * v[KVX] + v[KVY] * v[KVY] + v[KVZ] * v[KVZ];
      v2 = rtsafe(solver7, v2_old, 0.0, 1.0 - 1.e-11, 1.e-12, &ierr);
      if (ierr > 0) return 0;   /* the reference returns with errcode 0, v untouched */
      for (int k = 0; k < G.nv; k++) v[k] = uin[k];
      if (G.divclean) v[G.kglm] = u[G.kglm];
      w = w_solver;
      double rh = D * sqrt(1. - v2);
      double pg = gam * (w * (1. - v2) - rh);
      v[KRH] = rh;
      v[KPR] = fmax2(pg, eps);
    } else {
      double v2_old = v[KVX] * v[KVX] + v[KVY] * v[KVY] + v[KVZ] * v[KVZ] * g[2];
      v2 = rtsafe(solver8, v2_old, 0.0, 1.0 - 1.e-11, 1.e-12, &ierr);
      if ((v2 >= 1) || (ierr > 0)) return 778;
      for (int k = 0; k < G.nv; k++) v[k] = uin[k];
      if (G.divclean) v[G.kglm] = u[G.kglm];
      w = w_solver;
      double rh = D * sqrt(1. - v2);
      v[KPR] = 0.25 * (1 - v2) * w;
      v[KRH] = rh;
    }
    double vb = SB / w;
    v[KVX] = (sx + vb * bx) / (w + B2);
    v[KVY] = (sy + vb * by) / (w + B2);
    v[KVZ] = (sz / g[2] + vb * bz) / (w + B2);
    return 0;
  }
  if (G.algo_solver == 6) {
    /* exact quartic for e = 3p, system.f90:1000-1158 (Minkowski norms) */
    double v2_old = v[KVX] * v[KVX] + v[KVY] * v[KVY] + v[KVZ] * v[KVZ];
    double oldsol9 = 4 * v[KPR] / (1 - v2_old);
    for (int k = 0; k < G.nv; k++) uin[k] = gm * u[k];
    double DD = uin[KRH], sx = uin[KVX], sy = uin[KVY], sz = uin[KVZ], UU = uin[KPR];
    double bx = uin[kbx], by = uin[kby], bz = uin[kbz];
    double S2 = sx * sx + sy * sy + sz * sz;
    double B2 = bx * bx + by * by + bz * bz;
    double SB = (sx * bx + sy * by + sz * bz);
    double SB2 = SB * SB;
    double b9 = 8 * B2 - 4 * UU;
    double c9 = 7 * (B2 * B2) - 8 * B2 * UU + S2;
    double d9 = 2 * (B2 * B2 * B2) - 4 * (B2 * B2) * UU + 2 * B2 * S2;
    double e9 = SB2 * B2;
    double p9 = (24 * c9 - 3 * (b9 * b9)) / 72;
    double q9 = ((b9 * b9 * b9) - 12 * b9 * c9 + 72 * d9) / 216;
    double dnull9 = (c9 * c9) - 3 * b9 * d9 + 36 * e9;
    double deinz9 = 2 * (c9 * c9 * c9) - 9 * b9 * c9 * d9 + 27 * (b9 * b9) * e9 + 81 * (d9 * d9) - 216 * c9 * e9;
    double discr9 = (deinz9 * deinz9) - 4 * (dnull9 * dnull9 * dnull9);
    double sbig9 = 0, sol9 = -1, v2 = 0;
    int exact9 = 1;
    if (discr9 > 0) {
      double droot9 = sqrt(discr9);
      double qbig9 = pow(0.5 * (deinz9 + droot9), 1. / 3.);
      double drad9 = -2 * p9 / 3 + (qbig9 + dnull9 / qbig9) / 9;
      if (drad9 <= 0) return 12;
      sbig9 = 0.5 * sqrt(drad9);
    } else if (dnull9 > 0) {
      double phi9 = acos(deinz9 / (2 * sqrt(dnull9 * dnull9 * dnull9)));
      sbig9 = 0.5 * sqrt(-2. * p9 / 3 + 2. / 9. * sqrt(dnull9) * cos(phi9 / 3.));
    } else {
      if (solver10_approximate(b9, c9, d9, oldsol9, &sol9)) return 12;
      exact9 = 0;
      v2 = get_x9(sol9, S2, B2, SB2);
      if ((v2 < 0) || (v2 > 1) || (sol9 < 0)) return 12;
    }
    if (exact9) {
      double rad129 = -4 * (sbig9 * sbig9) - 2 * p9 + q9 / sbig9;
      double rad349 = -4 * (sbig9 * sbig9) - 2 * p9 - q9 / sbig9;
      double rapp9 = -b9 / 12;
      double sol1 = -1, sol2 = -1, sol3 = -1, sol4 = -1, dist9 = 1.e100, v9;
      if (rad129 >= 0) {
        double root129 = 0.5 * sqrt(rad129);
        sol1 = rapp9 - sbig9 - root129;
        sol2 = rapp9 - sbig9 + root129;
      }
      if (rad349 >= 0) {
        double root349 = 0.5 * sqrt(rad349);
        sol3 = rapp9 + sbig9 - root349;
        sol4 = rapp9 + sbig9 + root349;
      }
      sol9 = -1.;
      if (sol2 > 0) {
        v9 = get_x9(sol2, S2, B2, SB2);
        if ((v9 >= 0) && (v9 < 1)) {
          sol9 = sol2; v2 = v9; dist9 = fabs(oldsol9 - sol9);
          if (sol1 > 0) {
            v9 = get_x9(sol1, S2, B2, SB2);
            if ((v9 >= 0) && (v9 < 1)) {
              double d19 = fabs(oldsol9 - sol1);
              if (d19 <= dist9) { sol9 = sol1; dist9 = d19; v2 = v9; }
            }
          }
        }
      }
      if (sol4 > 0) {
        v9 = get_x9(sol4, S2, B2, SB2);
        if ((v9 >= 0) && (v9 < 1)) {
          double d49 = fabs(oldsol9 - sol4);
          if (d49 <= dist9) { sol9 = sol4; dist9 = d49; v2 = v9; }
        }
        if (sol3 > 0) {
          v9 = get_x9(sol3, S2, B2, SB2);
          if ((v9 >= 0) && (v9 < 1)) {
            double d39 = fabs(oldsol9 - sol3);
            if (d39 <= dist9) { sol9 = sol3; dist9 = d39; v2 = v9; }
          }
        }
      }
      if (sol9 < 0) return 12;
    }
    for (int k = 0; k < G.nv; k++) v[k] = uin[k];
    v[KRH] = DD * sqrt(1. - v2);
    v[KPR] = sol9 / 4 * (1 - v2);
    double vb = SB / sol9;
    v[KVX] = (sx + vb * bx) / (sol9 + B2);
    v[KVY] = (sy + vb * by) / (sol9 + B2);
    v[KVZ] = (sz + vb * bz) / (sol9 + B2);
    if (G.divclean) v[G.kglm] = u[G.kglm];
    return 0;
  }
  return 11;   /* algorithm 3 (LU-Newton) is not restated: getfun never sets Fnorm */
}

/* system_cons2prim, system.f90:423-665: ideal HD Newton on p (511-562) and
 * viscous nested iteration (564-665) */
int sys_cons2prim(double *u, double *v) {
  if (G.rmhd) return rmhd_cons2prim(u, v);
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
  /* pr_eos, dprdrh, dprden: when eos_pressure fails (out of table range)
   * it leaves its outputs unset and the Newton step below still uses them
   * (system.f90:534-538, 627-631: errcode is only checked after the loop).
   * Before the first successful call the reference reads uninitialised
   * locals there, whose values depend on the stack; this restatement (and
   * the GPU build) start them at 0 and count such steps in REP_C2P_UNDEF. */
  double pr, w = 0, v2, glf, rh = 0, en, drhdpr, dendpr, pr_eos = 0, dprdrh = 0, dprden = 0, f, df, dpr;
  int eos_set = 0;

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
      if (errcode == 0) eos_set = 1;
      else if (!eos_set) G.nrep[REP_C2P_UNDEF]++;
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
        if (errcode == 0) eos_set = 1;
        else if (!eos_set) G.nrep[REP_C2P_UNDEF]++;
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
  return mhd_cons2prim(u, v);
}

/* system_flux, system.f90:1316-1583 (HD and viscous branches) */
int sys_flux(double *v, double *u, double *f, double *vf, int dir) {
  if (G.rmhd) return rmhd_flux(v, u, f, vf, dir);
  double rh = v[KRH], vx = v[KVX], vy = v[KVY], vz = v[KVZ], pr = v[KPR], en;
  int err = eos_energy(rh, pr, &en);
  if (err > 0) return err;
  const double *g = G.g_cov;
  double v2 = g[0] * vx * vx + g[1] * vy * vy + g[2] * vz * vz;
  if (v2 > 1) return 333;
  double glf = 1. / sqrt(1. - v2);
  double w;
  double E[3] = {0, 0, 0}, EXB[3] = {0, 0, 0}, E2 = 0, B2 = 0, EEM = 0;
  if (!G.viscosity) {
    w = en + pr;
    if (G.mhd) {
      compute_E_ideal(&v[KVX], &v[G.kbx], E);
      E2 = E[0] * E[0] + E[1] * E[1] + E[2] * E[2] / g[2];
      system_EXB(E, &v[G.kbx], EXB);
      B2 = v[G.kbx] * v[G.kbx] + v[G.kby] * v[G.kby] + v[G.kbz] * v[G.kbz] * g[2];
      EEM = 0.5 * (E2 + B2);
    }
  } else {
    w = en + pr + v[KPIBU];
  }
  double ww = w * (glf * glf);
  double pitt = 0, pitx = 0, pity = 0, pitz = 0;
  u[KRH] = rh * glf;
  if (!G.viscosity && G.mhd) {
    u[KVX] = ww * vx + EXB[0];
    u[KVY] = ww * vy + EXB[1];
    u[KVZ] = ww * vz * g[2] + EXB[2];
    u[KPR] = ww - pr + EEM;
    u[G.kbx] = v[G.kbx]; u[G.kby] = v[G.kby]; u[G.kbz] = v[G.kbz];
    if (G.divclean) u[G.kglm] = v[G.kglm] / G.gp;
  } else if (G.viscosity) {
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
  const double Bx = G.mhd ? v[G.kbx] : 0, By = G.mhd ? v[G.kby] : 0, Bz = G.mhd ? v[G.kbz] : 0;
  if (dir == 1) {
    wwdir = ww * vx; vdir = vx; gcov_dir = g[0];
    f[KRH] = rh * glf * vx;
    if (!G.viscosity && G.mhd) {
      f[KVX] = wwdir * vx - E[0] * E[0] - Bx * Bx + pr + EEM;
      f[KVY] = wwdir * vy - E[0] * E[1] - Bx * By;
      f[KVZ] = wwdir * vz * g[2] - E[0] * E[2] - Bx * Bz * g[2];
      f[KPR] = wwdir + EXB[0];
      f[G.kbx] = 0.;
      f[G.kby] = -E[2] * G.gm;
      f[G.kbz] = +E[1] * G.gm;
      if (G.divclean) { f[G.kbx] = v[G.kglm]; f[G.kglm] = v[G.kbx] / G.gp; }
    } else if (G.viscosity) {
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
    if (!G.viscosity && G.mhd) {
      f[KVX] = wwdir * vx - E[1] * E[0] - By * Bx;
      f[KVY] = wwdir * vy - E[1] * E[1] - By * By + pr + EEM;
      f[KVZ] = wwdir * vz * g[2] - E[1] * E[2] - By * Bz * g[2];
      f[KPR] = wwdir + EXB[1];
      f[G.kbx] = E[2] * G.gm;
      f[G.kby] = 0.;
      f[G.kbz] = -E[0] * G.gm;
      if (G.divclean) { f[G.kby] = v[G.kglm]; f[G.kglm] = v[G.kby] / G.gp; }
    } else if (G.viscosity) {
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
    if (!G.viscosity && G.mhd) {
      f[KVX] = wwdir * vx - E[2] / g[2] * E[0] - Bz * Bx;
      f[KVY] = wwdir * vy - E[2] / g[2] * E[1] - Bz * By;
      f[KVZ] = wwdir * vz * g[2] - E[2] * E[2] / g[2] - Bz * Bz * g[2] + pr + EEM;
      f[KPR] = wwdir + EXB[2] / g[2];
      f[G.kbx] = -E[1] * G.gm;
      f[G.kby] = E[0] * G.gm;
      f[G.kbz] = 0.;
      if (G.divclean) { f[G.kbz] = v[G.kglm] / g[2]; f[G.kglm] = v[G.kbz] / G.gp; }
    } else if (G.viscosity) {
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
  if (G.mhd && !G.viscosity) {
    double bb2 = B2 - E2;
    ca2 = bb2 / (bb2 + (en + pr));
  }
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

/* The bulk viscosity: the reference's zeta = 2 eta (1/3 - c_s^2)
 * (viscous.f90:186-190), or, when set with orc_set_bulk_peak, a zeta/s
 * peaked at T_c, zeta = hbar zs_max exp(-(T - T_c)^2 / (2 w^2)) s (no
 * counterpart in the reference) */
double bulk_zeta(double eta, double cs2, double temp, double s) {
  if (!G.bulkvis) return 0.;
  if (G.zeta_peak) {
    const double x = (temp - G.zs_tc) / G.zs_w;
    return HBAR * G.zs_max * exp(-0.5 * x * x) * s;
  }
  return 2. * eta * (1. / 3. - cs2);
}

void orc_set_bulk_peak(double zs_max, double t_c, double width) {
  G.zeta_peak = 1;
  G.zs_max = zs_max;
  G.zs_tc = t_c;
  G.zs_w = width;
}

/* get_viscous_parameters, viscous.f90:172-199 */
static int get_viscous_parameters(double rh, double pr, double *eta, double *zeta, double *tau_pi, double *tau_pi_big) {
  double en, temp, s, cs2;
  int err = get_derived_data(rh, pr, &en, &temp, &s);
  int err2 = eos_sound(rh, en, pr, &cs2);
  if (err2) err = err2;
  *eta = HBAR * G.eta_over_s * s;
  if (G.init_type == SHEAR_VISCOUS_1D) *eta = G.eta_over_s;
  *zeta = bulk_zeta(*eta, cs2, temp, s);
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
    const double t2 = pow(t, 2.), t3 = t * t * t;   /* flang lowers t**3. to two multiplications */
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
  if (G.rmhd) return rmhd_source(ix, iy, iz, v, src, src_stiff);
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
      if (G.mhd) {
        double E[3];
        compute_E_ideal(&v[KVX], &v[G.kbx], E);
        double B2 = v[G.kbx] * v[G.kbx] + v[G.kby] * v[G.kby] + v[G.kbz] * v[G.kbz] * g[2];
        double E2 = E[0] * E[0] + E[1] * E[1] + E[2] * E[2] / g[2];
        double EEM = 0.5 * (E2 + B2);
        double e3 = E[2] / g[2];
        src[KPR] = -t * (ww * vz * vz - e3 * e3 - v[G.kbz] * v[G.kbz] + (EEM + pr) / g[2]);
      } else {
        src[KPR] = -t * (ww * vz * vz + pr / g[2]);
      }
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
