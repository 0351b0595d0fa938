/* ORACLE - test infrastructure, not product code (see oracle.h).
 * Reconstruction library: restates holib.f90 (reference).  Each routine
 * takes f with f[i] = value of cell i for i in [-ngc, n+ngc] (the pointer
 * points at element 0) and returns gl[i], gr[i], i = 0..n: the left and
 * right states at the interface i+1/2.
 */
#include "oracle_int.h"

/* Fortran SIGN(1.,d): +1 for d >= +0, -1 for negative d (incl. -0). */
static inline double sgn(double d) { return copysign(1.0, d); }

/* mm2, holib.f90:440-452 */
static double mm2(double d1, double d2) {
  double s1 = sgn(d1), s2 = sgn(d2);
  return .5 * (s1 + s2) * fmin2(fabs(d1), fabs(d2));
}

/* mm4, holib.f90:472-487 */
static double mm4(double d1, double d2, double d3, double d4) {
  double s1 = sgn(d1), s2 = sgn(d2), s3 = sgn(d3), s4 = sgn(d4);
  double m = fmin2(fmin2(fmin2(fabs(d1), fabs(d2)), fabs(d3)), fabs(d4));
  return .125 * (s1 + s2) * fabs((s1 + s3) * (s1 + s4)) * m;
}

/* ceno3, holib.f90:491-511 (alpha = 0.7, first minimum as MINLOC) */
static double ceno3(const double d[3]) {
  int is0 = (int)sgn(d[0]), is1 = (int)sgn(d[1]), is2 = (int)sgn(d[2]);
  int s = is0 + is1 + is2;
  if (s == 3 || s == -3) {
    double w[3] = {fabs(d[0]), 0.7 * fabs(d[1]), fabs(d[2])};
    int m = 0;
    if (w[1] < w[m]) m = 1;
    if (w[2] < w[m]) m = 2;
    return d[m];
  }
  return 0.;
}

/* mp5, holib.f90:515-536 */
static double mp5(double f, double f1, double f2, double f3, double f4, double f5) {
  double r = f;
  double ful = f3 + 2 * (f3 - f2);
  double fmp = f3 + mm2(f4 - f3, ful - f3);
  if ((f - f3) * (f - fmp) > 0.) {
    double dm = f1 - 2 * f2 + f3;
    double d0 = f2 - 2 * f3 + f4;
    double dp = f3 - 2 * f4 + f5;
    double fmd = .5 * (f3 + f4) - .5 * mm4(4 * d0 - dp, 4 * dp - d0, d0, dp);
    double flc = f3 + .5 * (f3 - f2) + (4. / 3.) * mm4(4 * d0 - dm, 4 * dm - d0, d0, dm);
    double fmin = fmax2(fmin2(fmin2(f3, f4), fmd), fmin2(fmin2(f3, ful), flc));
    double fmax = fmin2(fmax2(fmax2(f3, f4), fmd), fmax2(fmax2(f3, ful), flc));
    r = f + mm2(fmin - f, fmax - f);
  }
  return r;
}

/* holib_rectvd2, holib.f90:47-77 (mm2 limiter) */
static void rec_tvd2(const double *f, double *gl, double *gr, int n) {
  for (int i = 1; i <= n; i++) {
    double dp = .5 * mm2(f[i] - f[i - 1], f[i + 1] - f[i]);
    gl[i] = f[i] + dp;
    gr[i - 1] = f[i] - dp;
  }
  gl[0] = f[0] + .5 * mm2(f[0] - f[-1], f[1] - f[0]);
  gr[n] = f[n + 1] - .5 * mm2(f[n + 1] - f[n], f[n + 2] - f[n + 1]);
}

/* holib_recceno3, holib.f90:81-126 */
static void rec_ceno3(const double *f, double *gl, double *gr, int n) {
  const double c11 = 3 / 8., c21 = -10 / 8., c31 = 15 / 8.;
  const double c12 = -1 / 8., c22 = 6 / 8., c32 = 3 / 8.;
  const double c13 = 3 / 8., c23 = 6 / 8., c33 = -1 / 8.;
  for (int i = 0; i <= n; i++) {
    double f1 = f[i - 2], f2 = f[i - 1], f3 = f[i], f4 = f[i + 1], f5 = f[i + 2], d[3];
    double p = f3 + .5 * mm2(f3 - f2, f4 - f3);
    d[0] = c11 * f1 + c21 * f2 + c31 * f3 - p;
    d[1] = c12 * f2 + c22 * f3 + c32 * f4 - p;
    d[2] = c13 * f3 + c23 * f4 + c33 * f5 - p;
    gl[i] = p + ceno3(d);
    f1 = f[i + 3]; f2 = f[i + 2]; f3 = f[i + 1]; f4 = f[i]; f5 = f[i - 1];
    p = f3 + .5 * mm2(f3 - f2, f4 - f3);
    d[0] = c11 * f1 + c21 * f2 + c31 * f3 - p;
    d[1] = c12 * f2 + c22 * f3 + c32 * f4 - p;
    d[2] = c13 * f3 + c23 * f4 + c33 * f5 - p;
    gr[i] = p + ceno3(d);
  }
}

/* holib_recweno3, holib.f90:130-173 */
static void rec_weno3(const double *f, double *gl, double *gr, int n) {
  const double d1 = 1 / 3., d2 = 2 / 3., eps = 1.e-6;
  const double c11 = -1 / 2., c21 = 3 / 2., c12 = 1 / 2., c22 = 1 / 2.;
  for (int i = 0; i <= n; i++) {
    double f1 = f[i - 1], f2 = f[i], f3 = f[i + 1];
    double b1 = (f2 - f1) * (f2 - f1), b2 = (f3 - f2) * (f3 - f2);
    double a1 = d1 / ((b1 + eps) * (b1 + eps)), a2 = d2 / ((b2 + eps) * (b2 + eps));
    double p1 = c11 * f1 + c21 * f2, p2 = c12 * f2 + c22 * f3;
    gl[i] = (a1 * p1 + a2 * p2) / (a1 + a2);
    f1 = f[i + 2]; f2 = f[i + 1]; f3 = f[i];
    b1 = (f2 - f1) * (f2 - f1); b2 = (f3 - f2) * (f3 - f2);
    a1 = d1 / ((b1 + eps) * (b1 + eps)); a2 = d2 / ((b2 + eps) * (b2 + eps));
    p1 = c11 * f1 + c21 * f2; p2 = c12 * f2 + c22 * f3;
    gr[i] = (a1 * p1 + a2 * p2) / (a1 + a2);
  }
}

/* One WENO5 side from the 5-point stencil (f1..f5), holib.f90:206-217 */
static double weno5_side(double f1, double f2, double f3, double f4, double f5) {
  const double d1 = 1 / 16., d2 = 10 / 16., d3 = 5 / 16., eps = 1.e-6;
  const double c11 = 3 / 8., c21 = -10 / 8., c31 = 15 / 8.;
  const double c12 = -1 / 8., c22 = 6 / 8., c32 = 3 / 8.;
  const double c13 = 3 / 8., c23 = 6 / 8., c33 = -1 / 8.;
  double t, s;
  t = f3 - 2 * f2 + f1; s = 3. * f3 - 4. * f2 + f1;
  double beta1 = 13. * (t * t) + 3. * (s * s);
  t = f4 - 2 * f3 + f2; s = f4 - f2;
  double beta2 = 13. * (t * t) + 3. * (s * s);
  t = f5 - 2 * f4 + f3; s = f5 - 4. * f4 + 3. * f3;
  double beta3 = 13. * (t * t) + 3. * (s * s);
  double q1 = beta1 / 12. + eps, q2 = beta2 / 12. + eps, q3 = beta3 / 12. + eps;
  double alpha1 = d1 / (q1 * q1), alpha2 = d2 / (q2 * q2), alpha3 = d3 / (q3 * q3);
  double p1 = c11 * f1 + c21 * f2 + c31 * f3;
  double p2 = c12 * f2 + c22 * f3 + c32 * f4;
  double p3 = c13 * f3 + c23 * f4 + c33 * f5;
  return (alpha1 * p1 + alpha2 * p2 + alpha3 * p3) / (alpha1 + alpha2 + alpha3);
}

/* holib_recweno5, holib.f90:177-232 */
static void rec_weno5(const double *f, double *gl, double *gr, int n) {
  for (int i = 0; i <= n; i++) {
    gl[i] = weno5_side(f[i - 2], f[i - 1], f[i], f[i + 1], f[i + 2]);
    gr[i] = weno5_side(f[i + 3], f[i + 2], f[i + 1], f[i], f[i - 1]);
  }
}

/* holib_recppm4, holib.f90:236-268 (mm2 limiter) */
static void rec_ppm4(const double *f, double *gl, double *gr, int n) {
  for (int i = 0; i <= n; i++) {
    double f1 = f[i - 1], f2 = f[i], f3 = f[i + 1], f4 = f[i + 2];
    gl[i] = f2 + .5 * (f3 - f2) - .125 * (mm2(f4 - f3, f3 - f2) - mm2(f3 - f2, f2 - f1));
    f1 = f[i + 2]; f2 = f[i + 1]; f3 = f[i]; f4 = f[i - 1];
    gr[i] = f2 + .5 * (f3 - f2) - .125 * (mm2(f4 - f3, f3 - f2) - mm2(f3 - f2, f2 - f1));
  }
}

/* holib_recmpe3, holib.f90:272-293 */
static void rec_mpe3(const double *f, double *gl, double *gr, int n) {
  const double d1 = -1 / 8., d2 = 6 / 8., d3 = 3 / 8.;
  for (int i = 0; i <= n; i++) {
    gl[i] = d1 * f[i - 1] + d2 * f[i] + d3 * f[i + 1];
    gr[i] = d1 * f[i + 2] + d2 * f[i + 1] + d3 * f[i];
  }
  for (int i = 0; i <= n; i++) {
    gl[i] = mp5(gl[i], f[i - 2], f[i - 1], f[i], f[i + 1], f[i + 2]);
    gr[i] = mp5(gr[i], f[i + 3], f[i + 2], f[i + 1], f[i], f[i - 1]);
  }
}

/* holib_recmpe5, holib.f90:297-318 */
static void rec_mpe5(const double *f, double *gl, double *gr, int n) {
  const double d1 = 3 / 128., d2 = -20 / 128., d3 = 90 / 128., d4 = 60 / 128., d5 = -5 / 128.;
  for (int i = 0; i <= n; i++) {
    gl[i] = d1 * f[i - 2] + d2 * f[i - 1] + d3 * f[i] + d4 * f[i + 1] + d5 * f[i + 2];
    gr[i] = d1 * f[i + 3] + d2 * f[i + 2] + d3 * f[i + 1] + d4 * f[i] + d5 * f[i - 1];
  }
  for (int i = 0; i <= n; i++) {
    gl[i] = mp5(gl[i], f[i - 2], f[i - 1], f[i], f[i + 1], f[i + 2]);
    gr[i] = mp5(gr[i], f[i + 3], f[i + 2], f[i + 1], f[i], f[i - 1]);
  }
}

/* holib_recmpe7, holib.f90:322-348 */
static void rec_mpe7(const double *f, double *gl, double *gr, int n) {
  const double d1 = -5 / 1024., d2 = 42 / 1024., d3 = -175 / 1024., d4 = 700 / 1024.;
  const double d5 = 525 / 1024., d6 = -70 / 1024., d7 = 7 / 1024.;
  for (int i = 0; i <= n; i++) {
    gl[i] = d1 * f[i - 3] + d2 * f[i - 2] + d3 * f[i - 1] + d4 * f[i]
          + d5 * f[i + 1] + d6 * f[i + 2] + d7 * f[i + 3];
    gr[i] = d1 * f[i + 4] + d2 * f[i + 3] + d3 * f[i + 2] + d4 * f[i + 1]
          + d5 * f[i] + d6 * f[i - 1] + d7 * f[i - 2];
  }
  for (int i = 0; i <= n; i++) {
    gl[i] = mp5(gl[i], f[i - 2], f[i - 1], f[i], f[i + 1], f[i + 2]);
    gr[i] = mp5(gr[i], f[i + 3], f[i + 2], f[i + 1], f[i], f[i - 1]);
  }
}

/* ghost cells each method reads, work.f90:427-518 */
int rec_ngc(int code) {
  switch (code) {
    case 1: case 3: case 5: return 2;
    case 8: return 4;
    default: return 3;
  }
}

/* work_rec dispatch, work.f90:420-525; codes as oracle/ref_harness.f90 */
void rec_pencil(int code, const double *f, double *gl, double *gr, int n) {
  switch (code) {
    case 1: rec_tvd2(f, gl, gr, n); break;
    case 2: rec_ceno3(f, gl, gr, n); break;
    case 3: rec_weno3(f, gl, gr, n); break;
    case 4: rec_weno5(f, gl, gr, n); break;
    case 5: rec_ppm4(f, gl, gr, n); break;
    case 6: rec_mpe3(f, gl, gr, n); break;
    case 7: rec_mpe5(f, gl, gr, n); break;
    case 8: rec_mpe7(f, gl, gr, n); break;
    default: rec_weno5(f, gl, gr, n); break;
  }
}
