/* ORACLE - test infrastructure, not product code (see oracle.h).
 *
 * Resistive MHD, with and without second-order viscosity, on the grid update
 * of evolve.f90.  The reference has no working path for it: resistive MHD
 * forces algorithm 7 (common.f90:2108-2118), which system_cons2prim lacks
 * (system.f90:1160-1162), and viscosity together with MHD is refused
 * (common.f90:1904-1908).  This restatement is of the build's own design,
 * assembled from the reference's pieces:
 *  - layout of system_setup (system.f90:80-124): after the hydro (5) or
 *    viscous (12) variables come B^x, B^y, B^eta, E^x, E^y, E^eta and the
 *    Dedner scalar psi (divclean is required), nv = 12 or 19;
 *  - conserved variables and fluxes: the fluid part exactly as the hydro or
 *    viscous system (system.f90:164-263, 1316-1583), plus the
 *    electromagnetic part of the MHD branch with E evolved: E x B and
 *    (E^2 + B^2)/2 in S and in the energy, the Maxwell stresses in the
 *    momentum fluxes, the induction (B) and Ampere (E) fluxes, the GLM
 *    fluxes of psi and the normal B (system.f90:1425-1549);
 *  - sources (system.f90:1587-1842): the fluid sources as the hydro or
 *    viscous system, the electromagnetic stress in the Bjorken energy
 *    source, and Ohm's current in the E equations: -rho_c v explicitly and
 *    -gamma [sigma (E - (v.E) v + v x B) + sigma_chi (B - (v.B) v - v x E)]
 *    as the stiff part (system.f90:1653-1662), with rho_c = div E by
 *    central differences (system.f90:1901-1918);
 *  - con2prim: E and B are read off the conserved variables, their energy
 *    and momentum subtracted, and the hydro or viscous inversion run on the
 *    fluid part; inside an IMEX stage the implicit relation
 *    E = E* + alpha dt R(E, v) is solved with it by fixed-point iteration on
 *    v around a 3x3 linear solve for E (the role of the missing
 *    algorithm 7, whose locals system.f90:494-497 declare a 3x3 matrix).
 * Where the reference's pieces disagree with each other the metric-
 * consistent form is used: covariant E_eta = g33 E^eta in the fluxes as in
 * prim2cons (system.f90:1369 drops g33), v x B as in
 * system_compute_E_in_ideal_MHD (system_crossprod drops the metric),
 * v.B computed in Minkowski too (system.f90:1752-1753 computes v.E twice),
 * the evolved E in the Bjorken energy source (system.f90:1647 uses the
 * ideal-MHD E), the stiff source times sqrt(-g) like the other sources,
 * and conductivities in fm^-1 (sigma_el / hbar c, as the initial field of
 * glaubermc.f90:1014-1016 uses them).  The HLL speeds are those of light
 * for B, E and psi and the fast magnetosonic ones (with B^2 - E^2 in place
 * of B^2) for the fluid variables; the time step is limited by the speed
 * of light.
 */
#include <stdlib.h>
#include <string.h>
#include "oracle_int.h"

static const double HBARC = 0.197326;   /* common.f90:51 */

/* The grid's settings switched to the plain fluid system (hydro or
 * viscous) and back: the fluid part of every operator below is the
 * restatement of that system itself. */
typedef struct { int mhd, rmhd, divclean, nv, nu, nov, nou; } saved_sys;

static saved_sys fluid_mode(void) {
  saved_sys s = {G.mhd, G.rmhd, G.divclean, G.nv, G.nu, G.nov, G.nou};
  G.mhd = 0; G.rmhd = 0; G.divclean = 0;
  G.nv = G.nu = G.viscosity ? 12 : 5;
  G.nov = G.nou = 5;
  return s;
}

static void full_mode(saved_sys s) {
  G.mhd = s.mhd; G.rmhd = s.rmhd; G.divclean = s.divclean;
  G.nv = s.nv; G.nu = s.nu; G.nov = s.nov; G.nou = s.nou;
}

int rmhd_nfluid(void) { return G.viscosity ? 12 : 5; }

/* electromagnetic energy and momentum of contravariant E, B:
 * Ec covariant E, X = E x B (covariant, times sqrt(-g) as system_EXB),
 * E2, B2 */
static void em_terms(const double *E, const double *B, double *Ec, double *X, double *E2, double *B2) {
  const double *g = G.g_cov;
  Ec[0] = E[0]; Ec[1] = E[1]; Ec[2] = E[2] * g[2];
  *E2 = Ec[0] * E[0] + Ec[1] * E[1] + Ec[2] * E[2];
  *B2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2] * g[2];
  X[0] = G.gp * (E[1] * B[2] - E[2] * B[1]);
  X[1] = G.gp * (E[2] * B[0] - E[0] * B[2]);
  X[2] = G.gp * (E[0] * B[1] - E[1] * B[0]);
}

/* contravariant a x b for contravariant a, b (the -E of
 * system_compute_E_in_ideal_MHD, made contravariant) */
static void cross_contra(const double *a, const double *b, double *c) {
  c[0] = G.gp * (a[1] * b[2] - a[2] * b[1]);
  c[1] = G.gp * (a[2] * b[0] - a[0] * b[2]);
  c[2] = G.gp * (a[0] * b[1] - a[1] * b[0]) / G.g_cov[2];
}

static double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + G.g_cov[2] * a[2] * b[2];
}

int rmhd_prim2cons(double *v, double *u, int initial) {
  const int nf = rmhd_nfluid();
  saved_sys s = fluid_mode();
  int e = initial && G.viscosity ? sys_prim2cons_0(v, u) : sys_prim2cons(v, u);
  full_mode(s);
  if (e > 0) return e;
  const double *B = &v[G.kbx], *E = &v[G.kex];
  double Ec[3], X[3], E2, B2;
  em_terms(E, B, Ec, X, &E2, &B2);
  const double eem = 0.5 * (E2 + B2);
  const double gp = G.gp;
  u[KVX] = u[KVX] + gp * X[0];
  u[KVY] = u[KVY] + gp * X[1];
  u[KVZ] = u[KVZ] + gp * X[2];
  u[KPR] = u[KPR] + gp * eem;
  for (int k = 0; k < 3; k++) { u[G.kbx + k] = gp * B[k]; u[G.kex + k] = gp * E[k]; }
  u[G.kglm] = v[G.kglm];   /* (psi / gp) * gp, system.f90:253 */
  (void)nf;
  return 0;
}

/* fluid part of u for given E, B: S - sqrt(-g) E x B, energy - sqrt(-g) EEM */
static void fluid_part(const double *u, const double *E, const double *B, double *uf) {
  const int nf = rmhd_nfluid();
  double Ec[3], X[3], E2, B2;
  em_terms(E, B, Ec, X, &E2, &B2);
  const double gp = G.gp;
  for (int k = 0; k < nf; k++) uf[k] = u[k];
  uf[KVX] = u[KVX] - gp * X[0];
  uf[KVY] = u[KVY] - gp * X[1];
  uf[KVZ] = u[KVZ] - gp * X[2];
  uf[KPR] = u[KPR] - gp * (0.5 * (E2 + B2));
}

static int fluid_cons2prim(double *uf, double *v) {
  saved_sys s = fluid_mode();
  int e = sys_cons2prim(uf, v);
  full_mode(s);
  return e;
}

int rmhd_cons2prim(double *u, double *v) {
  double B[3], E[3], uf[MAXNV];
  for (int k = 0; k < 3; k++) { B[k] = u[G.kbx + k] * G.gm; E[k] = u[G.kex + k] * G.gm; }
  fluid_part(u, E, B, uf);
  int e = fluid_cons2prim(uf, v);
  for (int k = 0; k < 3; k++) { v[G.kbx + k] = B[k]; v[G.kex + k] = E[k]; }
  v[G.kglm] = u[G.kglm];
  return e;
}

/* The implicit part of an IMEX stage: u holds U*; finds E with
 * E = E* + adt R(E, v) where R is the stiff current (per unit sqrt(-g)) and
 * v the fluid velocity of the inversion with that E; B and the fluid's
 * conserved D, S, energy do not change.  Fixed point: solve the linear
 * system for E at the current v, invert the fluid with that E, repeat until
 * E changes by at most 1e-12 of the field scale max(|E|, |B|) (the last
 * inversion is then the one with the E before the final solve).  Every
 * fluid inversion starts from the primitives v held on entry, as the one
 * inversion of an explicit stage does: the viscous inversion stops its
 * velocity iteration at an absolute tolerance (system.f90:511-665), so
 * chaining the inversions would carry that iteration on and make the
 * result depend on where the loop stops.  Writes the new E into u. */
int rmhd_cons2prim_implicit(double *u, double *v, double adt) {
  const double sig = G.sigma_el / HBARC, sch = G.sigma_chi / HBARC;
  double B[3], Es[3], E[3], uf[MAXNV];
  for (int k = 0; k < 3; k++) { B[k] = u[G.kbx + k] * G.gm; Es[k] = u[G.kex + k] * G.gm; E[k] = v[G.kex + k]; }
  /* the starting guesses the fluid inversion reads: pressure and velocity */
  const double g0[4] = {v[KVX], v[KVY], v[KVZ], v[KPR]};
  int e = 0;
  for (int it = 0; it < 50; it++) {
    const double *vel = &v[KVX];
    const double v2 = G.g_cov[0] * vel[0] * vel[0] + G.g_cov[1] * vel[1] * vel[1] + G.g_cov[2] * vel[2] * vel[2];
    const double glf = 1. / sqrt(1. - v2);
    const double a = adt * glf * sig, b = adt * glf * sch;
    double vxB[3], vB = dot3(vel, B), vl[3] = {vel[0], vel[1], G.g_cov[2] * vel[2]};
    cross_contra(vel, B, vxB);
    double r[3], M[3][3];
    for (int i = 0; i < 3; i++) r[i] = Es[i] - a * vxB[i] - b * (B[i] - vB * vel[i]);
    /* (v x E)^i = C_ij E^j */
    const double gp = G.gp, g3 = G.g_cov[2];
    const double C[3][3] = {{0., -gp * vel[2], gp * vel[1]},
                            {gp * vel[2], 0., -gp * vel[0]},
                            {-gp * vel[1] / g3, gp * vel[0] / g3, 0.}};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) M[i][j] = (i == j ? 1. + a : 0.) - a * vel[i] * vl[j] - b * C[i][j];
    const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                       M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                       M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    double En[3];
    En[0] = (r[0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (r[1] * M[2][2] - M[1][2] * r[2]) +
             M[0][2] * (r[1] * M[2][1] - M[1][1] * r[2])) / det;
    En[1] = (M[0][0] * (r[1] * M[2][2] - M[1][2] * r[2]) - r[0] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
             M[0][2] * (M[1][0] * r[2] - r[1] * M[2][0])) / det;
    En[2] = (M[0][0] * (M[1][1] * r[2] - r[1] * M[2][1]) - M[0][1] * (M[1][0] * r[2] - r[1] * M[2][0]) +
             r[0] * (M[1][0] * M[2][1] - M[1][1] * M[2][0])) / det;
    /* the change of E against the field scale max(|E|, |B|), not against
     * |E| itself, which can be orders of magnitude smaller than B */
    double dE = 0., mE = 0.;
    for (int k = 0; k < 3; k++) {
      dE = fmax2(dE, fabs(En[k] - E[k]));
      mE = fmax2(mE, fmax2(fabs(En[k]), fabs(B[k])));
      E[k] = En[k];
    }
    /* converged: v is the inversion with the previous E, which equals E to
     * 1e-12 of the field scale */
    if (it > 0 && dE <= 1e-12 * mE) break;
    fluid_part(u, E, B, uf);
    v[KVX] = g0[0]; v[KVY] = g0[1]; v[KVZ] = g0[2]; v[KPR] = g0[3];
    e = fluid_cons2prim(uf, v);
    if (e) break;
  }
  for (int k = 0; k < 3; k++) { v[G.kbx + k] = B[k]; v[G.kex + k] = E[k]; u[G.kex + k] = E[k] * G.gp; }
  v[G.kglm] = u[G.kglm];
  return e;
}

int rmhd_flux(double *v, double *u, double *f, double *vf, int dir) {
  saved_sys s = fluid_mode();
  int e = sys_flux(v, u, f, vf, dir);
  full_mode(s);
  if (e > 0) return e;
  const double *g = G.g_cov;
  const double gp = G.gp, gm = G.gm;
  const double *B = &v[G.kbx], *E = &v[G.kex];
  double Ec[3], X[3], E2, B2;
  em_terms(E, B, Ec, X, &E2, &B2);
  const double eem = 0.5 * (E2 + B2);
  const double psi = v[G.kglm];
  double fm[3], fe, fb[3], fE[3], fpsi;
  /* Maxwell stress -E^d E_i - B^d B_i + delta_di EEM, Poynting flux, induction
   * and Ampere fluxes, GLM (system.f90:1431-1546) */
  if (dir == 1) {
    fm[0] = -E[0] * Ec[0] - B[0] * B[0] + eem;
    fm[1] = -E[0] * Ec[1] - B[0] * B[1];
    fm[2] = -E[0] * Ec[2] - B[0] * B[2] * g[2];
    fe = X[0];
    fb[0] = psi; fb[1] = -Ec[2] * gm; fb[2] = Ec[1] * gm;
    fE[0] = 0.; fE[1] = B[2] * g[2] * gm; fE[2] = -B[1] * g[1] * gm;
    fpsi = B[0] / gp;
  } else if (dir == 2) {
    fm[0] = -E[1] * Ec[0] - B[1] * B[0];
    fm[1] = -E[1] * Ec[1] - B[1] * B[1] + eem;
    fm[2] = -E[1] * Ec[2] - B[1] * B[2] * g[2];
    fe = X[1];
    fb[0] = Ec[2] * gm; fb[1] = psi; fb[2] = -Ec[0] * gm;
    fE[0] = -B[2] * g[2] * gm; fE[1] = 0.; fE[2] = B[0] * g[0] * gm;
    fpsi = B[1] / gp;
  } else {
    fm[0] = -E[2] * Ec[0] - B[2] * B[0];
    fm[1] = -E[2] * Ec[1] - B[2] * B[1];
    fm[2] = -E[2] * Ec[2] - B[2] * B[2] * g[2] + eem;
    fe = X[2] / g[2];
    fb[0] = -Ec[1] * gm; fb[1] = Ec[0] * gm; fb[2] = psi / g[2];
    fE[0] = B[1] * g[1] * gm; fE[1] = -B[0] * g[0] * gm; fE[2] = 0.;
    fpsi = B[2] / gp;
  }
  u[KVX] = u[KVX] + gp * X[0];
  u[KVY] = u[KVY] + gp * X[1];
  u[KVZ] = u[KVZ] + gp * X[2];
  u[KPR] = u[KPR] + gp * eem;
  f[KVX] = f[KVX] + gp * fm[0];
  f[KVY] = f[KVY] + gp * fm[1];
  f[KVZ] = f[KVZ] + gp * fm[2];
  f[KPR] = f[KPR] + gp * fe;
  for (int k = 0; k < 3; k++) {
    u[G.kbx + k] = gp * B[k]; u[G.kex + k] = gp * E[k];
    f[G.kbx + k] = gp * fb[k]; f[G.kex + k] = gp * fE[k];
  }
  u[G.kglm] = psi;
  f[G.kglm] = gp * fpsi;
  /* fast magnetosonic speeds of the fluid variables (system.f90:1557-1580)
   * with B^2 - E^2 for B^2 */
  const double rh = v[KRH], pr = v[KPR], vx = v[KVX], vy = v[KVY], vz = v[KVZ];
  double en, cs2;
  eos_energy(rh, pr, &en);
  e = eos_sound(rh, en, pr, &cs2);
  if (e > 0) return e;
  const double bb2 = fmax2(B2 - E2, 0.);
  const double ca2 = bb2 / (bb2 + (en + pr));
  const double v2 = g[0] * vx * vx + g[1] * vy * vy + g[2] * vz * vz;
  const double vdir = dir == 1 ? vx : dir == 2 ? vy : vz;
  const double gcov_dir = g[dir - 1];
  const double a2 = cs2 + ca2 - cs2 * ca2;
  const double den = 1. / (1. - v2 * a2);
  const double vf1 = den * (1. - a2) * vdir;
  const double vf2 = den * sqrt(a2 * (1. - v2) * ((1. - v2 * a2) / gcov_dir - (1. - a2) * (vdir * vdir)));
  vf[0] = vf1 + vf2;
  vf[1] = vf1 - vf2;
  return 0;
}

/* charge density rho_c = div E (central differences of the contravariant E,
 * system.f90:1901-1918; neighbours outside the grid by the boundary rule) */
static double div_e(int ix, int iy, int iz) {
  const size_t nc = ncell();
  const double *Ex = &G.v[(size_t)G.kex * nc], *Ey = &G.v[(size_t)(G.kex + 1) * nc], *Ez = &G.v[(size_t)(G.kex + 2) * nc];
  double r = 0.;
  int lo, hi;
#define NB(i, n, d) (G.boundary_cond == 0 ? (((i) + (d) + (n)) % (n)) : ((i) + (d) < 0 ? 0 : ((i) + (d) >= (n) ? (n) - 1 : (i) + (d))))
  if (G.nx > 1) {
    lo = NB(ix, G.nx, -1); hi = NB(ix, G.nx, 1);
    r = r + (Ex[cell_index(hi, iy, iz)] - Ex[cell_index(lo, iy, iz)]) * (G.ddx / 2.);
  }
  if (G.ny > 1) {
    lo = NB(iy, G.ny, -1); hi = NB(iy, G.ny, 1);
    r = r + (Ey[cell_index(ix, hi, iz)] - Ey[cell_index(ix, lo, iz)]) * (G.ddy / 2.);
  }
  if (G.nz > 1) {
    lo = NB(iz, G.nz, -1); hi = NB(iz, G.nz, 1);
    r = r + (Ez[cell_index(ix, iy, hi)] - Ez[cell_index(ix, iy, lo)]) * (G.ddz / 2.);
  }
#undef NB
  return r;
}

int rmhd_source(int ix, int iy, int iz, double *v, double *src, double *src_stiff) {
  for (int k = 0; k < G.nv; k++) { src[k] = 0.; src_stiff[k] = 0.; }
  saved_sys s = fluid_mode();
  int e = sys_source(ix, iy, iz, v, src, src_stiff);
  full_mode(s);
  if (e > 0) return e;
  const double *g = G.g_cov;
  const double gp = G.gp;
  const double *vel = &v[KVX], *B = &v[G.kbx], *E = &v[G.kex];
  if (G.coordinates == BJORKEN) {
    double Ec[3], X[3], E2, B2;
    em_terms(E, B, Ec, X, &E2, &B2);
    const double eem = 0.5 * (E2 + B2);
    src[KPR] = src[KPR] + (-G.t * (eem / g[2] - E[2] * E[2] - B[2] * B[2])) * gp;
  }
  const double rhoc = div_e(ix, iy, iz);
  double s3[3];
  rmhd_stiff(v, s3);
  for (int k = 0; k < 3; k++) {
    src[G.kex + k] = -rhoc * vel[k] * gp;
    src_stiff[G.kex + k] = s3[k];
  }
  return 0;
}

/* the stiff (Ohm) current of the E equations per cell, times sqrt(-g):
 * -gamma [sig (E - (v.E) v + v x B) + sch (B - (v.B) v - v x E)] */
void rmhd_stiff(const double *v, double *s3) {
  const double *g = G.g_cov;
  const double *vel = &v[KVX], *B = &v[G.kbx], *E = &v[G.kex];
  const double v2 = g[0] * vel[0] * vel[0] + g[1] * vel[1] * vel[1] + g[2] * vel[2] * vel[2];
  const double glf = 1. / sqrt(1. - v2);
  const double sig = G.sigma_el / HBARC, sch = G.sigma_chi / HBARC;
  double vxB[3], vxE[3];
  cross_contra(vel, B, vxB);
  cross_contra(vel, E, vxE);
  const double vE = dot3(vel, E), vB = dot3(vel, B);
  for (int k = 0; k < 3; k++)
    s3[k] = -glf * (sig * (E[k] - vE * vel[k] + vxB[k]) + sch * (B[k] - vB * vel[k] - vxE[k])) * G.gp;
}
