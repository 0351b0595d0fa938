/* ORACLE - test infrastructure, not product code (see oracle.h).
 * Equation of state: restates eos.f90 of the reference with numerically=0
 * (the default set at common.f90:652).
 *  eqstate 1: p = e/3 (eos_data/pressure_vs_rh_en1.def, energy_den1.def,
 *             part_der_pr_vs_rh_en1.def), T = 0.148909 e^(1/4)
 *             (eos_data/temperature_gubsertest.def, the file that
 *             temperature.def links to).
 *  eqstate 2: ideal gas p = (Gamma-1)(e-rho), Gamma = 4/3 (eos.f90:60,
 *             eos_data/ files ending in 2.def).
 *  eqstate 3: tabulated, strict-">" bisection + linear interpolation
 *             (eos.f90:399-701).
 */
#include <stdio.h>
#include <stdlib.h>
#include "oracle_int.h"

static const double GAMMA_AD = 4.0 / 3.0;   /* eos.f90:60 */
static const double HC3 = .00768351;         /* eos.f90:275 */

/* Bisection of eos.f90:430-442: returns indxl (1-based) such that the
 * bracket is [indxl, indxl+1], using "x > a(indx)" to move the low end. */
static int tab_bracket(const double *a, int n, double x) {
  int lo = 1, hi = n, mid = (lo + hi) / 2;
  while (lo + 1 != hi) {
    if (x > a[mid - 1]) { lo = mid; mid = (lo + hi) / 2; }
    else { hi = mid; mid = (lo + hi) / 2; }
  }
  return lo;
}

/* read_data_from_tabulated_file, eos.f90:265-334 */
int eos_load_table(const char *path) {
  FILE *fp = fopen(path, "r");
  if (!fp) return 1;
  int n = 0;
  if (fscanf(fp, "%d", &n) != 1 || n < 2) { fclose(fp); return 1; }
  free(G.tab_t); free(G.tab_e); free(G.tab_p); free(G.tab_c);
  G.tab_t = malloc(sizeof(double) * n); G.tab_e = malloc(sizeof(double) * n);
  G.tab_p = malloc(sizeof(double) * n); G.tab_c = malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++) {
    double T, e, p, c;
    if (fscanf(fp, "%lf %lf %lf %lf", &T, &e, &p, &c) != 4) { fclose(fp); return 1; }
    G.tab_t[i] = T;
    G.tab_e[i] = e / HC3 * pow(T, 4.0);
    G.tab_p[i] = p / HC3 * pow(T, 4.0);
    G.tab_c[i] = c;
  }
  fclose(fp);
  G.entries = n;
  return 0;
}

/* eos_pressure: eos_pressure_ana (eos.f90:888-925) / get_pressure
 * (eos.f90:399-464) */
int eos_pressure(double rh, double en, double *pr, double *dprdrh, double *dprden) {
  if (G.eqstate == 1) {
    *pr = en / 3.0;
    *dprden = 1. / 3.;
    *dprdrh = 0.;
    return 0;
  }
  if (G.eqstate == 2) {
    *pr = (GAMMA_AD - 1.) * (en - rh);
    *dprden = GAMMA_AD - 1.;
    *dprdrh = -GAMMA_AD + 1.;
    return 0;
  }
  const double *E = G.tab_e, *P = G.tab_p, *T = G.tab_t;
  int n = G.entries;
  if (en < E[0]) return 123;
  if (en > E[n - 1]) return 124;
  int l = tab_bracket(E, n, en), h = l + 1;
  double p2 = P[h - 1], p1 = P[l - 1], e2 = E[h - 1], e1 = E[l - 1];
  double t2 = T[h - 1], t1 = T[l - 1];
  *pr = p1 + (p2 - p1) * (en - e1) / (e2 - e1);
  double dpdt = (p2 - p1) / (t2 - t1);
  double dedt = (e2 - e1) / (t2 - t1);
  *dprdrh = 0;
  *dprden = dpdt / dedt;
  return 0;
}

/* eos_energy: eos_energy_ana (eos.f90:1108-1200) / get_energy
 * (eos.f90:468-523) */
int eos_energy(double rh, double pr, double *en) {
  if (G.eqstate == 1 || G.eqstate == 2) {
    if (pr <= G.przero) {                 /* check_on_pressure, eos.f90:1118-1123 */
      *en = G.enezero * pr / G.przero;
      return 0;
    }
    if (G.eqstate == 1) *en = 3. * pr;
    else *en = rh + pr / (GAMMA_AD - 1.);
    return 0;
  }
  const double *E = G.tab_e, *P = G.tab_p;
  int n = G.entries;
  if (pr < P[0]) return 125;
  if (pr > P[n - 1]) return 126;
  int l = tab_bracket(P, n, pr), h = l + 1;
  double p2 = P[h - 1], p1 = P[l - 1], e2 = E[h - 1], e1 = E[l - 1];
  *en = e1 + (e2 - e1) * (pr - p1) / (p2 - p1);
  return 0;
}

/* eos_sound: eos_sound_ana (eos.f90:929-953) / get_sound (eos.f90:590-641) */
int eos_sound(double rh, double en, double pr, double *cs2) {
  if (G.eqstate == 1 || G.eqstate == 2) {
    double dprden, dprdrh;
    if (G.eqstate == 1) { dprden = 1. / 3.; dprdrh = 0.; }
    else { dprden = GAMMA_AD - 1.; dprdrh = -GAMMA_AD + 1.; }
    double c;
    if (en + pr == 0) c = dprden;
    else c = dprden + (rh / (en + pr)) * dprdrh;
    *cs2 = c > 0.0 ? c : 0.0;
    return 0;
  }
  const double *E = G.tab_e, *C = G.tab_c;
  int n = G.entries;
  if (en < E[0]) return 131;
  if (en > E[n - 1]) return 132;
  int l = tab_bracket(E, n, en), h = l + 1;
  double s2 = C[h - 1], s1 = C[l - 1], e2 = E[h - 1], e1 = E[l - 1];
  *cs2 = s1 + (s2 - s1) * (en - e1) / (e2 - e1);
  return 0;
}

/* get_derived_data: get_derived_data_ana (eos.f90:957-969, entropy
 * 985-997) / get_derived_data_tab (eos.f90:645-701) */
int get_derived_data(double rh, double pr, double *en, double *temp, double *s) {
  if (G.eqstate == 1 || G.eqstate == 2) {
    int err = eos_energy(rh, pr, en);
    *temp = 0.148909 * pow(*en, 0.25);
    if (*temp == 0) *s = 0.;
    else *s = (*en + pr) / *temp;
    return err;
  }
  const double *E = G.tab_e, *P = G.tab_p, *T = G.tab_t;
  int n = G.entries;
  if (pr < P[0]) return 135;
  if (pr > P[n - 1]) return 136;
  int l = tab_bracket(P, n, pr), h = l + 1;
  double p2 = P[h - 1], p1 = P[l - 1], e2 = E[h - 1], e1 = E[l - 1];
  double t2 = T[h - 1], t1 = T[l - 1];
  double f = (pr - p1) / (p2 - p1);
  *temp = t1 + (t2 - t1) * f;
  *en = e1 + (e2 - e1) * f;
  *s = (*en + pr) / *temp;
  return 0;
}
