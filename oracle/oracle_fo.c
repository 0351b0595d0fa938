/*
 * ORACLE - test infrastructure, not product code.
 *
 * Freeze-out hypersurface finder, restated for checking the GPU finder:
 *   - the set-up: allocate_hypersurface_arrays (hypersurface.f90:93-168) and
 *     the header values of get_hysufile_ready (1137-1200);
 *   - one slice, out_hyp (out.f90:1568-1645): the freeze-out field (T for
 *     freeze_type 0, e for 1, from get_derived_data) of the current state is
 *     shifted in; the first slice is only stored, every later one is
 *     compared with the stored one by find_hypersurface_{ideal,visco,mhd}
 *     (hypersurface.f90:172-229, 466-521, 775-830);
 *   - one element per crossing of freeze_value along tau (same cell, two
 *     slices) and along x, y, eta (neighbouring cells of the older slice),
 *     in the reference's loop order (ix outer, iz inner; tau, x, y, eta per
 *     cell), interpolated by set_fo_{ide,vis,mhd} (354-460, 645-769,
 *     984-1117) and signed by save_hypersurface_* (280-350, 568-641,
 *     907-980);
 *   - the printed row of each element and the text file hypersurface.txt
 *     (print_hypersurface_*, 231-276, 523-566, 832-905): 13 values for
 *     ideal hydro, + 7 stresses + pi^{tau mu} (get_derived_pi(_zz)_gcv3,
 *     viscous.f90:203-240, with g33 = tau_fo^2) + 1/(T^2 (e+p)) for viscous
 *     hydro, + B^x, B^y, B^eta for MHD.
 * Not restated: the derivative output (derivatives_out), the comoving
 * charge density of MHD (print_rho_comov) and the restart path.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "oracle_int.h"

#define FO_SUPERLUMINAL 1005   /* set_fo_*: "you have to fix this", exit(1) */

static struct {
  int on, ftype;
  double fv, t1, t2, temp_f, eng_f;
  double *f1, *f2;   /* field(1,:,:,:), field(2,:,:,:) */
  double *vb, *pr;   /* v_backup, primitives */
  double *rec;
  size_t n, cap;
} F;

void fo_release(void) {
  free(F.f1); free(F.f2); free(F.vb); free(F.pr); free(F.rec);
  memset(&F, 0, sizeof F);
}

int orc_fo_width(void) { return G.mhd ? 16 : (G.viscosity ? 25 : 13); }

/* hdr: (temp_f, eng_f) of get_hysufile_ready.  The reference passes an
 * unset density to the EoS there; the tabulated and p = e/3 EoS ignore it,
 * and 1 is used here. */
int orc_fo_setup(int ftype, double fv, double *hdr) {
  fo_release();
  if (ftype != 0 && ftype != 1) return -1;
  const size_t nc = ncell();
  F.f1 = calloc(nc, sizeof(double));
  F.f2 = calloc(nc, sizeof(double));
  F.vb = calloc(nc * G.nv, sizeof(double));
  F.pr = calloc(nc * G.nv, sizeof(double));
  F.ftype = ftype;
  F.fv = fv;
  F.on = 1;
  double prex, d1, d2, s;
  int e = 0;
  if (ftype == 0) {
    F.temp_f = fv;
    find_pressure_from_temperature(F.temp_f, &prex);
    e = eos_energy(1., prex, &F.eng_f);
  } else {
    F.eng_f = fv;
    e = eos_pressure(1., F.eng_f, &prex, &d1, &d2);
    if (e <= 0) e = get_derived_data(1., prex, &F.eng_f, &F.temp_f, &s);
  }
  hdr[0] = F.temp_f;
  hdr[1] = F.eng_f;
  return e > 0 ? e : 0;
}

/* get_derived_pi_gcv3 (viscous.f90:203-215); g_cov(1) = g_cov(2) = 1 */
static void pi_gcv3(const double *p, double gcv3, double *o) {
  const double g1 = G.g_cov[0], g2 = G.g_cov[1];
  o[1] = p[KPIXX] * p[KVX] * g1 + p[KPIXY] * p[KVY] * g2 + p[KPIXZ] * p[KVZ] * gcv3;
  o[2] = p[KPIXY] * p[KVX] * g1 + p[KPIYY] * p[KVY] * g2 + p[KPIYZ] * p[KVZ] * gcv3;
  o[3] = p[KPIXZ] * p[KVX] * g1 + p[KPIYZ] * p[KVY] * g2 + p[KPIZZ] * p[KVZ] * gcv3;
  o[0] = o[1] * p[KVX] * g1 + o[2] * p[KVY] * g2 + o[3] * p[KVZ] * gcv3;
}

/* get_derived_pi_zz_gcv3 (viscous.f90:219-240) */
static void pi_zz_gcv3(const double *p, double gcv3, double *o) {
  const double g1 = G.g_cov[0], g2 = G.g_cov[1];
  const double vx = p[KVX], vy = p[KVY], vz = p[KVZ];
  const double xx = p[KPIXX], yy = p[KPIYY], xy = p[KPIXY], xz = p[KPIXZ], yz = p[KPIYZ];
  o[0] = 1. / (1. - vz * vz * gcv3) * (vx * g1 * xy * vy * g2 + vx * vx * g1 * g1 * xx +
         vy * g1 * yz * vz * gcv3 + 2. * vx * g1 * xz * vz * gcv3 + vy * g1 * g1 * xy * vx +
         vy * vy * g1 * yy * g2 + vz * gcv3 * yz * vy * g2 - xx * g1 * vz * vz * gcv3 -
         yy * g2 * vz * vz * gcv3);
  o[1] = xx * vx * g1 + xy * vy * g2 + xz * vz * gcv3;
  o[2] = xy * vx * g1 + yy * vy * g2 + yz * vz * gcv3;
  o[3] = 1. / (1. - vz * vz * gcv3) * (xz * vx * g1 + xz * vx * g1 * vz * vz * gcv3 + yz * vy * g2 +
         vz * vx * g1 * xy * vy * g2 - vz * xx * g1 - vz * yy * g2 + vz * vx * vx * g1 * g1 * xx +
         vy * g1 * yz * vz * vz * gcv3 + vz * vy * g1 * g1 * xy * vx + vz * vy * vy * g1 * yy * g2);
}

/* set_fo_* + save_hypersurface_* + the derived part of print_hypersurface_*
 * for the crossing along direction d (0 tau, 1 x, 2 y, 3 eta) from cell
 * (ix, iy, iz) of the older slice. */
static int element(int d, int ix, int iy, int iz, double *r) {
  const size_t nc = ncell();
  const int jx = ix + (d == 1), jy = iy + (d == 2), jz = iz + (d == 3);
  const size_t c = cell_index(ix, iy, iz), cc = cell_index(jx, jy, jz);
  const double fa = F.f1[c], fb = d == 0 ? F.f2[c] : F.f1[cc];
  const double ta = F.t1, tb = d == 0 ? F.t2 : F.t1;
  const double *A = F.vb, *B = d == 0 ? F.pr : F.vb;
  const size_t cb = d == 0 ? c : cc;
  const double si = (fa - fb) >= 0. ? 1. : -1.;
  const double weight = (fa - F.fv) / (fa - fb);
#define INTERP(k) (A[(size_t)(k) * nc + c] + weight * (B[(size_t)(k) * nc + cb] - A[(size_t)(k) * nc + c]))
  double fo[13];
  fo[0] = ta + weight * (tb - ta);
  fo[1] = G.xc[ix] + weight * (G.xc[jx] - G.xc[ix]);
  fo[2] = G.yc[iy] + weight * (G.yc[jy] - G.yc[iy]);
  fo[3] = G.zc[iz] + weight * (G.zc[jz] - G.zc[iz]);
  fo[4] = INTERP(KRH);
  fo[5] = INTERP(KVX);
  fo[6] = INTERP(KVY);
  fo[7] = INTERP(KVZ);
  fo[8] = INTERP(KPR);
  const double betasq = fo[5] * fo[5] * G.g_cov[0] + fo[6] * fo[6] * G.g_cov[1] + fo[7] * fo[7] * G.g_cov[2];
  if (betasq > 1.0) return FO_SUPERLUMINAL;
  const double dt = fabs(F.t2 - F.t1);
  const double dx = G.dx, dy = G.dy, dz = G.nz > 1 ? G.dz : 1.0;
  fo[9] = (double)(d == 0) * dx * dy * dz * fo[0];
  fo[10] = (double)(jx - ix) * dt * dy * dz * fo[0];
  fo[11] = (double)(jy - iy) * dx * dt * dz * fo[0];
  fo[12] = (double)(jz - iz) * dx * dy * dt * fo[0];
  for (int i = 0; i < 9; i++) r[i] = fo[i];
  for (int i = 9; i < 13; i++) r[i] = si * fo[i];
  if (G.mhd) {
    r[13] = INTERP(G.kbx);
    r[14] = INTERP(G.kby);
    r[15] = INTERP(G.kbz);
  } else if (G.viscosity) {
    const int ks[7] = {KPIBU, KPIXY, KPIXZ, KPIYZ, KPIXX, KPIYY, KPIZZ};
    for (int i = 0; i < 7; i++) r[13 + i] = INTERP(ks[i]);
    /* print_hypersurface_visco: the EoS error flag is not checked there */
    double eng = 0., temp = 0., s = 0.;
    get_derived_data(r[4], r[8], &eng, &temp, &s);
    double fake[MAXNV] = {0};
    fake[KVX] = r[5]; fake[KVY] = r[6]; fake[KVZ] = r[7];
    fake[KPIXX] = r[17]; fake[KPIYY] = r[18]; fake[KPIZZ] = r[19];
    fake[KPIXY] = r[14]; fake[KPIXZ] = r[15]; fake[KPIYZ] = r[16];
    const double gcv3 = r[0] * r[0];
    if (G.obtained_zz) pi_zz_gcv3(fake, gcv3, r + 20);
    else pi_gcv3(fake, gcv3, r + 20);
    r[24] = 1.0 / (temp * temp * (eng + r[8]));
  }
#undef INTERP
  return 0;
}

static int push(int d, int ix, int iy, int iz) {
  const int w = orc_fo_width();
  if (F.n == F.cap) {
    F.cap = F.cap ? 2 * F.cap : 1024;
    F.rec = realloc(F.rec, sizeof(double) * w * F.cap);
  }
  int e = element(d, ix, iy, iz, F.rec + F.n * w);
  if (e) return e;
  F.n++;
  return 0;
}

static int crosses(double a, double b) { return (b - F.fv) * (F.fv - a) > 0.0; }

/* out_hyp: returns 0, an EoS error code (the reference sets run_crashed;
 * here nothing is changed) or FO_SUPERLUMINAL. */
int orc_fo_slice(long *nout) {
  *nout = 0;
  if (!F.on) return -1;
  const size_t nc = ncell();
  double *fnew = malloc(sizeof(double) * nc);
  for (size_t c = 0; c < nc; c++) {
    double en, temp, s;
    int e = get_derived_data(G.v[KRH * nc + c], G.v[KPR * nc + c], &en, &temp, &s);
    if (e > 0) { free(fnew); return e; }
    fnew[c] = F.ftype == 0 ? temp : en;
  }
  const int first = F.t2 == 0.;
  if (first) F.t2 = G.t;
  double *old = F.f1;
  F.f1 = F.f2;
  F.f2 = fnew;
  free(old);
  F.n = 0;
  if (first) {
    memcpy(F.pr, G.v, sizeof(double) * nc * G.nv);
    return 0;
  }
  memcpy(F.vb, F.pr, sizeof(double) * nc * G.nv);
  memcpy(F.pr, G.v, sizeof(double) * nc * G.nv);
  F.t1 = F.t2;
  F.t2 = G.t;
  for (int ix = 0; ix < G.nx; ix++)
    for (int iy = 0; iy < G.ny; iy++)
      for (int iz = 0; iz < G.nz; iz++) {
        const size_t c = cell_index(ix, iy, iz);
        int e = 0;
        if (crosses(F.f1[c], F.f2[c])) e = push(0, ix, iy, iz);
        if (!e && ix < G.nx - 1 && crosses(F.f1[c], F.f1[c + 1])) e = push(1, ix, iy, iz);
        if (!e && iy < G.ny - 1 && crosses(F.f1[c], F.f1[c + G.nx])) e = push(2, ix, iy, iz);
        if (!e && iz < G.nz - 1 && crosses(F.f1[c], F.f1[c + (size_t)G.nx * G.ny])) e = push(3, ix, iy, iz);
        if (e) return e;
      }
  *nout = (long)F.n;
  return 0;
}

void orc_fo_get(double *out) {
  memcpy(out, F.rec, sizeof(double) * F.n * orc_fo_width());
}

/* Fortran ES14.7 */
static void es14_7(char *b, double x) {
  char t[40];
  snprintf(t, sizeof t, "%.7E", x);
  char *e = strchr(t, 'E');
  if (e && strlen(e + 2) > 2) memmove(e, e + 1, strlen(e + 1) + 1);  /* +ddd: no E */
  snprintf(b, 40, "%14s", t);
}

/* list-directed output of a real(8) by the flang runtime: the shortest
 * round-trip digits, F form for 0.095 <= |x| < 1e15 (no leading zero),
 * E form (at least two exponent digits) otherwise */
static void list_real(char *b, double x) {
  if (x == 0.) { strcpy(b, "0."); return; }
  char t[40];
  int p;
  for (p = 1; p <= 17; p++) {
    snprintf(t, sizeof t, "%.*e", p - 1, x);
    if (strtod(t, NULL) == x) break;
  }
  char dig[24];
  int nd = 0;
  const char *q = t;
  if (*q == '-') q++;
  for (; *q && *q != 'e'; q++)
    if (*q != '.') dig[nd++] = *q;
  dig[nd] = 0;
  const int ex = atoi(q + 1);
  char *o = b;
  if (x < 0) *o++ = '-';
  const int f = (ex >= -1 && ex < 15) || (ex == -2 && dig[0] == '9' && nd > 1 && dig[1] >= '5');
  if (f) {
    if (ex >= 0) {
      for (int i = 0; i <= ex; i++) *o++ = i < nd ? dig[i] : '0';
      *o++ = '.';
      for (int i = ex + 1; i < nd; i++) *o++ = dig[i];
    } else {
      *o++ = '.';
      for (int i = 0; i < -ex - 1; i++) *o++ = '0';
      for (int i = 0; i < nd; i++) *o++ = dig[i];
    }
    *o = 0;
  } else {
    *o++ = dig[0];
    *o++ = '.';
    for (int i = 1; i < nd; i++) *o++ = dig[i];
    sprintf(o, "E%c%02d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
  }
}

/* header = 1: get_hysufile_ready's file with its header line (list-directed
 * freeze_type, temperature, energy density); header = 0: append the last
 * slice as print_hypersurface_* does ("(i6)" count, then one row per
 * element of ES14.7 values separated by two blanks). */
int orc_fo_write(const char *path, int header) {
  FILE *f = fopen(path, header ? "w" : "a");
  if (!f) return -2;
  if (header) {
    char a[40], b[40];
    list_real(a, F.temp_f);
    list_real(b, F.eng_f);
    fprintf(f, " %d %s %s\n", F.ftype, a, b);
  } else {
    const int w = orc_fo_width();
    if (F.n > 999999) fprintf(f, "******\n");
    else fprintf(f, "%6ld\n", (long)F.n);
    char v[40];
    for (size_t i = 0; i < F.n; i++) {
      for (int j = 0; j < w; j++) {
        es14_7(v, F.rec[i * w + j]);
        fputs(v, f);
        if (j < w - 1) fputs("  ", f);
      }
      fputc('\n', f);
    }
  }
  fclose(f);
  return 0;
}

/* the freeze-out part of out_dump (out.f90:1375-1380): thyp, time_arr,
 * v_backup (the older slice), field(2,:,:,:) */
int fo_dump_part(FILE *f, double thyp) {
  if (!F.on) return 0;
  const size_t nc = ncell();
  const double ta[2] = {F.t1, F.t2};
  int bad = fwrite(&thyp, sizeof(double), 1, f) != 1;
  bad |= fwrite(ta, sizeof(double), 2, f) != 2;
  bad |= fwrite(F.vb, sizeof(double), nc * G.nv, f) != nc * G.nv;
  bad |= fwrite(F.f2, sizeof(double), nc, f) != nc;
  return bad;
}

/* the freeze-out part of out_restart (out.f90:1440-1445) */
int fo_restart_part(FILE *f, double *thyp) {
  if (!F.on) return 0;
  const size_t nc = ncell();
  double ta[2];
  int bad = fread(thyp, sizeof(double), 1, f) != 1;
  bad |= fread(ta, sizeof(double), 2, f) != 2;
  bad |= fread(F.vb, sizeof(double), nc * G.nv, f) != nc * G.nv;
  bad |= fread(F.f2, sizeof(double), nc, f) != nc;
  F.t1 = ta[0];
  F.t2 = ta[1];
  return bad;
}
