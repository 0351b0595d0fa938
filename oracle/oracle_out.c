/*
 * ORACLE - test infrastructure, not product code.
 *
 * Output frames of out_vars (out.f90:232-591), restated for checking the GPU
 * writer: a Fortran stream file holding t, then the selected fields in the
 * order of type out_type (common.f90:122-151), each as an nx*ny*nz array
 * (x fastest), in FP64 (precision 8) or FP32 (precision 4).  Derived fields:
 * energy density, temperature, entropy density from get_derived_data;
 * pi^{t mu} (and pi^{eta eta} with TRACE_IMP 'zz') from get_derived_pi(_zz);
 * v0 = gamma; the ideal-MHD electric field of system_compute_E_in_ideal_MHD
 * (system.f90:1885-1897) with E_z made contravariant, and on resistive grids
 * the evolved E as held (out.f90:404-413).  The reference prints
 * pi^{eta eta} from the 'zz' work array, filled only when a pi^{t mu} field
 * is selected too; the same selection is required here.  The comoving
 * density (rc), the derivatives and the flows are not written.
 */
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>

#include "oracle.h"
#include "oracle_int.h"

enum { O_RH, O_VX, O_VY, O_VZ, O_PR, O_EN, O_TEMP, O_ENTR, O_BULK, O_PITT, O_PITX, O_PITY, O_PITZ,
       O_PIXY, O_PIXZ, O_PIYZ, O_PIXX, O_PIYY, O_PIZZ, O_V0, O_BX, O_BY, O_BZ, O_EX, O_EY, O_EZ,
       O_GLM, O_RC, O_N };

static int put(FILE *f, const double *a, size_t n, int prec) {
  if (prec == 8) return fwrite(a, sizeof(double), n, f) == n ? 0 : 1;
  for (size_t i = 0; i < n; i++) {
    float x = (float)a[i];
    if (fwrite(&x, sizeof(float), 1, f) != 1) return 1;
  }
  return 0;
}

/* sel: 28 flags; path: the frame file; tpath: time.dat (appended). */
int orc_out_frame(const int *sel, int prec, int iout, const char *path, const char *tpath) {
  const size_t n = ncell();
  if (sel[O_RC]) return -1;
  if (G.viscosity && sel[O_PIZZ] && !(sel[O_PITT] || sel[O_PITX] || sel[O_PITY] || sel[O_PITZ]))
    return -1;
  if (G.viscosity && sel[O_PIZZ] && !G.obtained_zz) return -1;
  FILE *f = fopen(path, "wb");
  if (!f) return -2;
  int err = 0;
  if (prec == 8) err |= fwrite(&G.t, sizeof(double), 1, f) != 1;
  else {
    float t4 = (float)G.t;
    err |= fwrite(&t4, sizeof(float), 1, f) != 1;
  }
  FILE *tf = fopen(tpath, "a");
  if (tf) {
    fprintf(tf, "%5d%16.9f\n", iout, G.t);
    fclose(tf);
  }
  double *a = malloc(sizeof(double) * n * 5);
  const int prims[5] = {KRH, KVX, KVY, KVZ, KPR};
  for (int q = 0; q < 5; q++)
    if (sel[q]) err |= put(f, G.v + (size_t)prims[q] * n, n, prec);
  if (sel[O_EN] || sel[O_TEMP] || sel[O_ENTR]) {
    for (size_t c = 0; c < n; c++) {
      double en, temp, s;
      int e = get_derived_data(G.v[KRH * n + c], G.v[KPR * n + c], &en, &temp, &s);
      if (e > 0) { free(a); fclose(f); return e; }
      a[c] = en; a[n + c] = temp; a[2 * n + c] = s;
    }
    for (int q = 0; q < 3; q++)
      if (sel[O_EN + q]) err |= put(f, a + (size_t)q * n, n, prec);
  }
  if (G.viscosity) {
    if (sel[O_BULK]) err |= put(f, G.v + (size_t)KPIBU * n, n, prec);
    if (sel[O_PITT] || sel[O_PITX] || sel[O_PITY] || sel[O_PITZ]) {
      double vv[MAXNV];
      for (size_t c = 0; c < n; c++) {
        for (int k = 0; k < G.nv; k++) vv[k] = G.v[k * n + c];
        if (G.obtained_zz) get_derived_pi_zz(vv, &a[c], &a[n + c], &a[2 * n + c], &a[3 * n + c], &a[4 * n + c]);
        else get_derived_pi(vv, &a[c], &a[n + c], &a[2 * n + c], &a[3 * n + c]);
      }
    }
    for (int q = 0; q < 4; q++)
      if (sel[O_PITT + q]) err |= put(f, a + (size_t)q * n, n, prec);
    const int pis[5] = {KPIXY, KPIXZ, KPIYZ, KPIXX, KPIYY};
    for (int q = 0; q < 5; q++)
      if (sel[O_PIXY + q]) err |= put(f, G.v + (size_t)pis[q] * n, n, prec);
    if (sel[O_PIZZ]) err |= put(f, a + 4 * n, n, prec);
  }
  if (sel[O_V0]) {
    const double *g = G.g_cov;
    for (size_t c = 0; c < n; c++) {
      double vx = G.v[KVX * n + c], vy = G.v[KVY * n + c], vz = G.v[KVZ * n + c];
      a[c] = 1. / sqrt(1. - (vx * vx) * g[0] - (vy * vy) * g[1] - (vz * vz) * g[2]);
    }
    err |= put(f, a, n, prec);
  }
  if (G.mhd) {
    if (sel[O_BX]) err |= put(f, G.v + (size_t)G.kbx * n, n, prec);
    if (sel[O_BY]) err |= put(f, G.v + (size_t)G.kby * n, n, prec);
    if (sel[O_BZ]) err |= put(f, G.v + (size_t)G.kbz * n, n, prec);
    if (G.rmhd) {
      /* out.f90:404-413: the evolved E as it is held */
      const int ke[3] = {G.kex, G.key, G.kez};
      for (int q = 0; q < 3; q++)
        if (sel[O_EX + q]) err |= put(f, G.v + (size_t)ke[q] * n, n, prec);
    } else if (sel[O_EX] || sel[O_EY] || sel[O_EZ]) {
      for (size_t c = 0; c < n; c++) {
        double V3[3] = {G.v[KVX * n + c], G.v[KVY * n + c], G.v[KVZ * n + c]};
        double B3[3] = {G.v[G.kbx * n + c], G.v[G.kby * n + c], G.v[G.kbz * n + c]};
        a[c] = -G.gp * (V3[1] * B3[2] - V3[2] * B3[1]);
        a[n + c] = -G.gp * (V3[2] * B3[0] - V3[0] * B3[2]);
        a[2 * n + c] = -G.gp * (V3[0] * B3[1] - V3[1] * B3[0]) / G.g_cov[2];
      }
      for (int q = 0; q < 3; q++)
        if (sel[O_EX + q]) err |= put(f, a + (size_t)q * n, n, prec);
    }
    if (G.divclean && sel[O_GLM]) err |= put(f, G.v + (size_t)G.kglm * n, n, prec);
  }
  free(a);
  fclose(f);
  return err ? -3 : 0;
}

/* out_dump (out.f90:1313-1390), serial path: dump.dat as a Fortran stream
 * file -- iout (4-byte integer), t, timeold, timeinterval, v, u (nv
 * fields each), vold, uold (krh..kvold_end), g_cov_3_old, and with the
 * freeze-out hypersurface on: thyp, time_arr, v_backup, field(2,:,:,:). */
int orc_dump(const char *path, int iout, double thyp) {
  FILE *f = fopen(path, "wb");
  if (!f) return -2;
  const size_t n = ncell();
  const int32_t io = iout;
  int bad = fwrite(&io, 4, 1, f) != 1;
  bad |= fwrite(&G.t, sizeof(double), 1, f) != 1;
  bad |= fwrite(&G.timeold, sizeof(double), 1, f) != 1;
  bad |= fwrite(&G.timeinterval, sizeof(double), 1, f) != 1;
  bad |= fwrite(G.v, sizeof(double), n * G.nv, f) != n * G.nv;
  bad |= fwrite(G.u, sizeof(double), n * G.nv, f) != n * G.nv;
  bad |= fwrite(G.vold, sizeof(double), n * G.kvold_end, f) != n * G.kvold_end;
  bad |= fwrite(G.uold, sizeof(double), n * G.kvold_end, f) != n * G.kvold_end;
  bad |= fwrite(&G.g_cov_3_old, sizeof(double), 1, f) != 1;
  bad |= fo_dump_part(f, thyp);
  fclose(f);
  return bad ? -3 : 0;
}

/* out_restart (out.f90:1394-1483) followed by system_metric and the tout
 * update of echo.f90:164-172 */
int orc_restart(const char *path, int *iout, double *thyp) {
  FILE *f = fopen(path, "rb");
  if (!f) return -2;
  const size_t n = ncell();
  int32_t io = 0;
  int bad = fread(&io, 4, 1, f) != 1;
  bad |= fread(&G.t, sizeof(double), 1, f) != 1;
  bad |= fread(&G.timeold, sizeof(double), 1, f) != 1;
  bad |= fread(&G.timeinterval, sizeof(double), 1, f) != 1;
  bad |= fread(G.v, sizeof(double), n * G.nv, f) != n * G.nv;
  bad |= fread(G.u, sizeof(double), n * G.nv, f) != n * G.nv;
  bad |= fread(G.vold, sizeof(double), n * G.kvold_end, f) != n * G.kvold_end;
  bad |= fread(G.uold, sizeof(double), n * G.kvold_end, f) != n * G.kvold_end;
  bad |= fread(&G.g_cov_3_old, sizeof(double), 1, f) != 1;
  bad |= fo_restart_part(f, thyp);
  fclose(f);
  *iout = io;
  system_metric();
  G.tout = G.t + G.dtout;
  return bad ? -3 : 0;
}
