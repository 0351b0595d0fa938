// Device physics for the ECHO-QGP grid update on gfx950: equation of state,
// reconstruction, fluxes, con2prim and sources, one cell or one interface per
// thread, FP64 throughout.  Operation order follows the reference; the
// library is built with -ffp-contract=fast, so multiply-adds become FMAs and
// the FP64 results agree with the reference to rounding, not bit for bit.
// Reference file:line citations on each function.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace echo {

// primitive / conserved slots, common.f90:63 (0-based)
enum { KRH = 0, KVX = 1, KVY = 2, KVZ = 3, KPR = 4, KPIBU = 5, KPIXY = 6,
       KPIXZ = 7, KPIYZ = 8, KPIXX = 9, KPIYY = 10, KPIZZ = 11 };

enum { MINKOWSKI = 1, BJORKEN = 2 };
enum { SHEAR_VISCOUS_1D = 2, GUBSER_VISCOUS = 5 };

// Evolved systems (system_setup, system.f90:67-128): hydro, viscous hydro,
// ideal MHD, ideal MHD with Dedner GLM divergence cleaning, and resistive
// MHD (E evolved, with the Dedner scalar) without and with viscosity.
enum { SYS_HD = 0, SYS_VISC = 1, SYS_MHD = 2, SYS_MHDGLM = 3, SYS_RMHD = 4, SYS_VRMHD = 5 };
// magnetic field and GLM slots of the ideal MHD systems (kbx..kbz, kglm)
enum { KBX = 5, KBY = 6, KBZ = 7, KGLM = 8 };
template <int SYS>
struct SysT {
  static constexpr bool RES = SYS == SYS_RMHD || SYS == SYS_VRMHD;
  static constexpr bool VISC = SYS == SYS_VISC || SYS == SYS_VRMHD;
  static constexpr bool MHD = SYS == SYS_MHD || SYS == SYS_MHDGLM;   // ideal MHD
  static constexpr bool GLM = SYS == SYS_MHDGLM || RES;
  // a resistive system is its fluid system (hydro or viscous, NF variables)
  // followed by B^x, B^y, B^eta, E^x, E^y, E^eta and psi (the layout of
  // system_setup, system.f90:80-124)
  static constexpr int FLUID = VISC ? SYS_VISC : SYS_HD;
  static constexpr int NF = VISC ? 12 : 5;
  static constexpr int NV = RES ? NF + 7 : (SYS == SYS_HD ? 5 : (SYS == SYS_VISC ? 12 : (SYS == SYS_MHD ? 8 : 9)));
  static constexpr int KB = RES ? NF : KBX;        // B^x
  static constexpr int KE = NF + 3;                // E^x (resistive systems)
  static constexpr int KG = RES ? NF + 6 : KGLM;   // psi
};

// Reconstruction codes (same numbering as the C-ABI echo_config.recal).
enum { REC_TVD2 = 1, REC_CENO3 = 2, REC_WENO3 = 3, REC_WENO5 = 4, REC_PPM4 = 5,
       REC_MPE3 = 6, REC_MPE5 = 7 };

struct Eos {
  int eqstate;          // 1 p=e/3, 2 ideal gas, 3 table
  int n;                // table entries
  double przero, enezero;
  const double* __restrict__ tt;   // temperature
  const double* __restrict__ te;   // energy density
  const double* __restrict__ tp;   // pressure
  const double* __restrict__ tc;   // c_s^2
  // bracket accelerator (see tab_bracket_fast); null pointers select the
  // plain bisection
  const int* __restrict__ ixe;     // bin starts over te
  const int* __restrict__ ixp;     // bin starts over tp
  int base_e, nb_e, base_p, nb_p;
};

struct Metric {
  double g1, g2, g3, gp, gm, g0, t;
};

__host__ __device__ __forceinline__ Metric make_metric(int coordinates, double t) {
  // system_metric_minkowski / _bjorken, system.f90:132-160
  Metric m;
  m.g1 = 1.; m.g2 = 1.; m.g0 = -1.; m.t = t;
  if (coordinates == BJORKEN) { m.g3 = t * t; m.gp = t; m.gm = 1. / t; }
  else { m.g3 = 1.; m.gp = 1.; m.gm = 1.; }
  return m;
}

__host__ __device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }

// ---------------------------------------------------------------------------
// Equation of state (eos.f90, numerically = 0)
// ---------------------------------------------------------------------------
constexpr double GAMMA_AD = 4.0 / 3.0;   // eos.f90:60

// strict-">" bisection of eos.f90:430-442 (1-based result indxl)
__host__ __device__ __forceinline__ int tab_bracket(const double* __restrict__ a, int n, double x) {
  int lo = 1, hi = n, mid = (lo + hi) >> 1;
  while (lo + 1 != hi) {
    if (x > a[mid - 1]) lo = mid; else hi = mid;
    mid = (lo + hi) >> 1;
  }
  return lo;
}

// Bracket accelerator.  For a strictly increasing table a[0..n-1] and
// a[0] <= x <= a[n-1], the bisection above returns max(1, c) with
// c = #{j : a[j] < x}, clamped to n-1, whatever path it takes.  c is found
// here by a short lower-bound search inside one bin of a monotone bin
// function: the bit pattern of the float nearest to x, shifted right by
// TAB_SHIFT (2048 bins per octave).  Because the bin function is
// monotone, entries in lower bins are < x and entries in higher bins are
// > x, so only the entries of x's own bin are searched: at most 2 for
// qcdIEOS0.dat, whose bin starts take 350 KB per column (any float range
// needs at most 2 MB).  The chain of dependent table loads is then the bin
// start and one or two entries (at 256 bins per octave, up to 13 entries
// and four search steps: 2% slower flux sweeps).  The result is the same
// index as tab_bracket.
constexpr int TAB_SHIFT = 12;

__host__ __device__ __forceinline__ int tab_bin(double x, int base, int nb) {
  float f = (float)x;
  unsigned u;
#ifdef __HIP_DEVICE_COMPILE__
  u = __float_as_uint(f);
#else
  __builtin_memcpy(&u, &f, 4);
#endif
  if (u >> 31) return 0;   // negative values, -0 included
  int b = (int)(u >> TAB_SHIFT) - base;
  return b < 0 ? 0 : (b >= nb ? nb - 1 : b);
}

__host__ __device__ __forceinline__ int tab_bracket_fast(const double* __restrict__ a,
                                                         const int* __restrict__ start, int base,
                                                         int nb, int n, double x) {
  int b = tab_bin(x, base, nb);
  int lo = start[b], hi = start[b + 1];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo < 1 ? 1 : (lo > n - 1 ? n - 1 : lo);
}

__host__ __device__ __forceinline__ int bracket_e(const Eos& E, double x) {
  return E.ixe ? tab_bracket_fast(E.te, E.ixe, E.base_e, E.nb_e, E.n, x) : tab_bracket(E.te, E.n, x);
}
__host__ __device__ __forceinline__ int bracket_p(const Eos& E, double x) {
  return E.ixp ? tab_bracket_fast(E.tp, E.ixp, E.base_p, E.nb_p, E.n, x) : tab_bracket(E.tp, E.n, x);
}

// eos_pressure_ana (eos.f90:888-925) / get_pressure (eos.f90:399-464)
__host__ __device__ __forceinline__ int eos_pressure(const Eos& E, double rh, double en, double& pr,
                                            double& dprdrh, double& dprden) {
  if (E.eqstate == 1) { pr = en / 3.0; dprden = 1. / 3.; dprdrh = 0.; return 0; }
  if (E.eqstate == 2) {
    pr = (GAMMA_AD - 1.) * (en - rh); dprden = GAMMA_AD - 1.; dprdrh = -GAMMA_AD + 1.;
    return 0;
  }
  if (en < E.te[0]) return 123;
  if (en > E.te[E.n - 1]) return 124;
  int l = bracket_e(E, en);
  double p2 = E.tp[l], p1 = E.tp[l - 1], e2 = E.te[l], e1 = E.te[l - 1];
  double t2 = E.tt[l], t1 = E.tt[l - 1];
  pr = p1 + (p2 - p1) * (en - e1) / (e2 - e1);
  double dpdt = (p2 - p1) / (t2 - t1);
  double dedt = (e2 - e1) / (t2 - t1);
  dprdrh = 0;
  dprden = dpdt / dedt;
  return 0;
}

// eos_energy_ana (eos.f90:1108-1200) / get_energy (eos.f90:468-523)
__host__ __device__ __forceinline__ int eos_energy(const Eos& E, double rh, double pr, double& en) {
  if (E.eqstate != 3) {
    if (pr <= E.przero) { en = E.enezero * pr / E.przero; return 0; }
    en = (E.eqstate == 1) ? 3. * pr : rh + pr / (GAMMA_AD - 1.);
    return 0;
  }
  if (pr < E.tp[0]) return 125;
  if (pr > E.tp[E.n - 1]) return 126;
  int l = bracket_p(E, pr);
  double p2 = E.tp[l], p1 = E.tp[l - 1], e2 = E.te[l], e1 = E.te[l - 1];
  en = e1 + (e2 - e1) * (pr - p1) / (p2 - p1);
  return 0;
}

// eos_sound_ana (eos.f90:929-953) / get_sound (eos.f90:590-641)
__host__ __device__ __forceinline__ int eos_sound(const Eos& E, double rh, double en, double pr, double& cs2) {
  if (E.eqstate != 3) {
    double dprden, dprdrh;
    if (E.eqstate == 1) { dprden = 1. / 3.; dprdrh = 0.; }
    else { dprden = GAMMA_AD - 1.; dprdrh = -GAMMA_AD + 1.; }
    double c = (en + pr == 0) ? dprden : dprden + (rh / (en + pr)) * dprdrh;
    cs2 = c > 0.0 ? c : 0.0;
    return 0;
  }
  if (en < E.te[0]) return 131;
  if (en > E.te[E.n - 1]) return 132;
  int l = bracket_e(E, en);
  double s2 = E.tc[l], s1 = E.tc[l - 1], e2 = E.te[l], e1 = E.te[l - 1];
  cs2 = s1 + (s2 - s1) * (en - e1) / (e2 - e1);
  return 0;
}

// eos_energy, then eos_sound at the energy it returns (the pair side_state
// evaluates).  For the table, en lies between te[l-1] and te[l] of the
// pressure bracket l; strictly inside, l is also the energy bracket (the
// count of te[j] < en), so that search is skipped; on a node it runs as in
// eos_sound.  Returns the energy error; the sound error (returned by
// eos_sound at the same point) goes to err_sound.
__host__ __device__ __forceinline__ int eos_energy_sound(const Eos& E, double rh, double pr, double& en,
                                                         double& cs2, int& err_sound) {
  err_sound = 0;
  if (E.eqstate != 3) {
    const int err = eos_energy(E, rh, pr, en);
    if (err > 0) return err;
    err_sound = eos_sound(E, rh, en, pr, cs2);
    return err;
  }
  if (pr < E.tp[0]) return 125;
  if (pr > E.tp[E.n - 1]) return 126;
  const int l = bracket_p(E, pr);
  const double p2 = E.tp[l], p1 = E.tp[l - 1], e2 = E.te[l], e1 = E.te[l - 1];
  en = e1 + (e2 - e1) * (pr - p1) / (p2 - p1);
  if (en < E.te[0]) { err_sound = 131; return 0; }
  if (en > E.te[E.n - 1]) { err_sound = 132; return 0; }
  const int m = (en > e1 && en < e2) ? l : bracket_e(E, en);
  const double s2 = E.tc[m], s1 = E.tc[m - 1], f2 = E.te[m], f1 = E.te[m - 1];
  cs2 = s1 + (s2 - s1) * (en - f1) / (f2 - f1);
  return 0;
}

// get_derived_data_ana (eos.f90:957-969) / get_derived_data_tab (645-701)
__host__ __device__ __forceinline__ int get_derived_data(const Eos& E, double rh, double pr, double& en,
                                                double& temp, double& s) {
  if (E.eqstate != 3) {
    int err = eos_energy(E, rh, pr, en);
    temp = 0.148909 * pow(en, 0.25);
    s = (temp == 0) ? 0. : (en + pr) / temp;
    return err;
  }
  if (pr < E.tp[0]) return 135;
  if (pr > E.tp[E.n - 1]) return 136;
  int l = bracket_p(E, pr);
  double p2 = E.tp[l], p1 = E.tp[l - 1], e2 = E.te[l], e1 = E.te[l - 1];
  double t2 = E.tt[l], t1 = E.tt[l - 1];
  double f = (pr - p1) / (p2 - p1);
  temp = t1 + (t2 - t1) * f;
  en = e1 + (e2 - e1) * f;
  s = (en + pr) / temp;
  return 0;
}

// 1/x for finite x > 0 in the normal range: the hardware reciprocal and two
// Newton-Raphson steps, within about an ulp of the correctly rounded value,
// in five instructions instead of the ten of the IEEE division sequence.
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

// 1/sqrt(x) for 0 < x <= 1 (Lorentz factors of subluminal states): the
// hardware reciprocal square root and two Newton-Raphson steps, within about
// an ulp of 1/sqrt(x); x = 0 gives +inf as the IEEE form does.
__device__ __forceinline__ double rsqrt_nr(double x) {
  double r = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  double e = __builtin_fma(-h * r, r, 0.5);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-h * r, r, 0.5);
  r = __builtin_fma(r, e, r);
  return x == 0.0 ? __builtin_inf() : r;
}

// ---------------------------------------------------------------------------
// Reconstruction (holib.f90).  f[0..5] = cells i-2 .. i+3 around the
// interface i+1/2; returns the left (gl) and right (gr) states there.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sgn1(double d) { return copysign(1.0, d); }

// mm2, holib.f90:440-452
__device__ __forceinline__ double mm2(double d1, double d2) {
  return .5 * (sgn1(d1) + sgn1(d2)) * dmin(fabs(d1), fabs(d2));
}
// mm4, holib.f90:472-487
__device__ __forceinline__ double mm4(double d1, double d2, double d3, double d4) {
  double s1 = sgn1(d1), s2 = sgn1(d2), s3 = sgn1(d3), s4 = sgn1(d4);
  double m = dmin(dmin(dmin(fabs(d1), fabs(d2)), fabs(d3)), fabs(d4));
  return .125 * (s1 + s2) * fabs((s1 + s3) * (s1 + s4)) * m;
}
// ceno3, holib.f90:491-511
__device__ __forceinline__ double ceno3(double d0, double d1, double d2) {
  int s = (int)sgn1(d0) + (int)sgn1(d1) + (int)sgn1(d2);
  if (s == 3 || s == -3) {
    double w0 = fabs(d0), w1 = 0.7 * fabs(d1), w2 = fabs(d2);
    double r = d0, wm = w0;
    if (w1 < wm) { wm = w1; r = d1; }
    if (w2 < wm) { r = d2; }
    return r;
  }
  return 0.;
}
// mp5, holib.f90:515-536
__device__ __forceinline__ double mp5(double f, double f1, double f2, double f3, double f4, double f5) {
  double r = f;
  double ful = f3 + 2 * (f3 - f2);
  double fmp = f3 + mm2(f4 - f3, ful - f3);
  if ((f - f3) * (f - fmp) > 0.) {
    double dm = f1 - 2 * f2 + f3;
    double d0 = f2 - 2 * f3 + f4;
    double dp = f3 - 2 * f4 + f5;
    double fmd = .5 * (f3 + f4) - .5 * mm4(4 * d0 - dp, 4 * dp - d0, d0, dp);
    double flc = f3 + .5 * (f3 - f2) + (4. / 3.) * mm4(4 * d0 - dm, 4 * dm - d0, d0, dm);
    double fmin = dmax(dmin(dmin(f3, f4), fmd), dmin(dmin(f3, ful), flc));
    double fmax = dmin(dmax(dmax(f3, f4), fmd), dmax(dmax(f3, ful), flc));
    r = f + mm2(fmin - f, fmax - f);
  }
  return r;
}

// One WENO5 side (holib.f90:206-217).  The reference's weights
// alpha_k = d_k / (beta_k/12 + eps)^2, normalised by their sum, are
// evaluated as  w_k = d_k * prod_{j != k} Q_j  with Q_j = (beta_j + 12 eps)^2:
// the same ratios (the factor 144 cancels), with one reciprocal per side
// instead of seven divisions (the sum of the weights is positive, at least
// (12 eps)^4 / 16, and finite).  Results agree with the reference form to
// rounding.
__device__ __forceinline__ double weno5_side(double f1, double f2, double f3, double f4, double f5) {
  const double d1 = 1 / 16., d2 = 10 / 16., d3 = 5 / 16., eps12 = 12. * 1.e-6;
  double t, s;
  t = f3 - 2 * f2 + f1; s = 3. * f3 - 4. * f2 + f1;
  double q1 = 13. * (t * t) + 3. * (s * s) + eps12;
  t = f4 - 2 * f3 + f2; s = f4 - f2;
  double q2 = 13. * (t * t) + 3. * (s * s) + eps12;
  t = f5 - 2 * f4 + f3; s = f5 - 4. * f4 + 3. * f3;
  double q3 = 13. * (t * t) + 3. * (s * s) + eps12;
  double Q1 = q1 * q1, Q2 = q2 * q2, Q3 = q3 * q3;
  double w1 = d1 * (Q2 * Q3), w2 = d2 * (Q1 * Q3), w3 = d3 * (Q1 * Q2);
  double p1 = 3 / 8. * f1 + -10 / 8. * f2 + 15 / 8. * f3;
  double p2 = -1 / 8. * f2 + 6 / 8. * f3 + 3 / 8. * f4;
  double p3 = 3 / 8. * f3 + 6 / 8. * f4 + -1 / 8. * f5;
  return (w1 * p1 + w2 * p2 + w3 * p3) * rcp_nr(w1 + w2 + w3);
}

__device__ __forceinline__ double weno3_side(double f1, double f2, double f3) {
  // holib.f90:148-158
  const double eps = 1.e-6;
  double b1 = (f2 - f1) * (f2 - f1), b2 = (f3 - f2) * (f3 - f2);
  double a1 = (1 / 3.) / ((b1 + eps) * (b1 + eps)), a2 = (2 / 3.) / ((b2 + eps) * (b2 + eps));
  double p1 = -1 / 2. * f1 + 3 / 2. * f2, p2 = 1 / 2. * f2 + 1 / 2. * f3;
  return (a1 * p1 + a2 * p2) / (a1 + a2);
}

__device__ __forceinline__ double ceno3_side(double f1, double f2, double f3, double f4, double f5) {
  // holib.f90:98-108
  double p = f3 + .5 * mm2(f3 - f2, f4 - f3);
  double d0 = 3 / 8. * f1 + -10 / 8. * f2 + 15 / 8. * f3 - p;
  double d1 = -1 / 8. * f2 + 6 / 8. * f3 + 3 / 8. * f4 - p;
  double d2 = 3 / 8. * f3 + 6 / 8. * f4 + -1 / 8. * f5 - p;
  return p + ceno3(d0, d1, d2);
}

template <int REC>
__device__ __forceinline__ void rec_face(const double* f, double& gl, double& gr) {
  // f[j] = cell (i-2+j)
  if constexpr (REC == REC_WENO5) {           // holib.f90:177-232
    gl = weno5_side(f[0], f[1], f[2], f[3], f[4]);
    gr = weno5_side(f[5], f[4], f[3], f[2], f[1]);
  } else if constexpr (REC == REC_PPM4) {     // holib.f90:236-268
    gl = f[2] + .5 * (f[3] - f[2]) - .125 * (mm2(f[4] - f[3], f[3] - f[2]) - mm2(f[3] - f[2], f[2] - f[1]));
    gr = f[3] + .5 * (f[2] - f[3]) - .125 * (mm2(f[1] - f[2], f[2] - f[3]) - mm2(f[2] - f[3], f[3] - f[4]));
  } else if constexpr (REC == REC_TVD2) {     // holib.f90:47-77
    gl = f[2] + .5 * mm2(f[2] - f[1], f[3] - f[2]);
    gr = f[3] - .5 * mm2(f[3] - f[2], f[4] - f[3]);
  } else if constexpr (REC == REC_MPE5) {     // holib.f90:297-318
    const double d1 = 3 / 128., d2 = -20 / 128., d3 = 90 / 128., d4 = 60 / 128., d5 = -5 / 128.;
    double l = d1 * f[0] + d2 * f[1] + d3 * f[2] + d4 * f[3] + d5 * f[4];
    double r = d1 * f[5] + d2 * f[4] + d3 * f[3] + d4 * f[2] + d5 * f[1];
    gl = mp5(l, f[0], f[1], f[2], f[3], f[4]);
    gr = mp5(r, f[5], f[4], f[3], f[2], f[1]);
  } else if constexpr (REC == REC_WENO3) {    // holib.f90:130-173
    gl = weno3_side(f[1], f[2], f[3]);
    gr = weno3_side(f[4], f[3], f[2]);
  } else if constexpr (REC == REC_CENO3) {    // holib.f90:81-126
    gl = ceno3_side(f[0], f[1], f[2], f[3], f[4]);
    gr = ceno3_side(f[5], f[4], f[3], f[2], f[1]);
  } else {                                    // REC_MPE3, holib.f90:272-293
    const double d1 = -1 / 8., d2 = 6 / 8., d3 = 3 / 8.;
    double l = d1 * f[1] + d2 * f[2] + d3 * f[3];
    double r = d1 * f[4] + d2 * f[3] + d3 * f[2];
    gl = mp5(l, f[0], f[1], f[2], f[3], f[4]);
    gr = mp5(r, f[5], f[4], f[3], f[2], f[1]);
  }
}

// ---------------------------------------------------------------------------
// Viscous helpers (viscous.f90)
// ---------------------------------------------------------------------------
struct PiDer { double pitt, pitx, pity, pitz; };

// get_derived_pi (227-239) and get_derived_pi_zz (243-279); for 'zz' the
// traceless pi^zz is written back into p[KPIZZ] as the reference does.
template <bool ZZ>
__device__ __forceinline__ PiDer derived_pi(double* p, const Metric& M) {
  const double g1 = M.g1, g2 = M.g2, g3 = M.g3;
  const double vx = p[KVX], vy = p[KVY], vz = p[KVZ];
  const double xx = p[KPIXX], yy = p[KPIYY], xy = p[KPIXY], xz = p[KPIXZ], yz = p[KPIYZ];
  PiDer r;
  if constexpr (!ZZ) {
    r.pitx = xx * vx * g1 + xy * vy * g2 + xz * vz * g3;
    r.pity = xy * vx * g1 + yy * vy * g2 + yz * vz * g3;
    r.pitz = xz * vx * g1 + yz * vy * g2 + p[KPIZZ] * vz * g3;
    r.pitt = r.pitx * vx * g1 + r.pity * vy * g2 + r.pitz * vz * g3;
  } else {
    // 1/(1 - v_eta^2 g33) once, by reciprocal (positive for subluminal v)
    const double ivz = rcp_nr(1. - vz * vz * g3);
    r.pitt = ivz * (vx * g1 * xy * vy * g2 + vx * vx * g1 * g1 * xx +
             vy * g1 * yz * vz * g3 + 2. * vx * g1 * xz * vz * g3 + vy * g1 * g1 * xy *
             vx + vy * vy * g1 * yy * g2 + vz * g3 * yz * vy * g2 - xx * g1 * vz * vz *
             g3 - yy * g2 * vz * vz * g3);
    r.pitx = xx * vx * g1 + xy * vy * g2 + xz * vz * g3;
    r.pity = xy * vx * g1 + yy * vy * g2 + yz * vz * g3;
    r.pitz = ivz * (xz * vx * g1 + xz * vx * g1 * vz * vz * g3 + yz * vy *
             g2 + vz * vx * g1 * xy * vy * g2 - vz * xx * g1 - vz * yy * g2 + vz * vx *
             vx * g1 * g1 * xx + vy * g1 * yz * vz * vz * g3 + vz * vy * g1 * g1 *
             xy * vx + vz * vy * vy * g1 * yy * g2);
    p[KPIZZ] = (ivz * rcp_nr(g3)) * (vx * g1 * xy * vy * g2 - xx * g1 - yy * g2 +
             vx * vx * g1 * g1 * xx + vy * g1 * yz * vz * g3 + 2. * vx * g1 * xz * vz *
             g3 + vy * g1 * g1 * xy * vx + vy * vy * g1 * yy * g2 + vz * g3 * yz *
             vy * g2);
  }
  return r;
}

// ---------------------------------------------------------------------------
// prim2cons, system.f90:164-263
// ---------------------------------------------------------------------------
// system_compute_E_in_ideal_MHD (system.f90:1882-1897): covariant E = -v x B
// times sqrt(-g); system_EXB (1846-1862): E x B
__device__ __forceinline__ void ideal_E(const Metric& M, const double* v, double* E) {
  E[0] = -M.gp * (v[KVY] * v[KBZ] - v[KVZ] * v[KBY]);
  E[1] = -M.gp * (v[KVZ] * v[KBX] - v[KVX] * v[KBZ]);
  E[2] = -M.gp * (v[KVX] * v[KBY] - v[KVY] * v[KBX]);
}
__device__ __forceinline__ void e_cross_b(const Metric& M, const double* E, const double* v, double* X) {
  double c0 = E[0], c1 = E[1], c2 = E[2] / M.g3;
  X[0] = M.gp * (c1 * v[KBZ] - c2 * v[KBY]);
  X[1] = M.gp * (c2 * v[KBX] - c0 * v[KBZ]);
  X[2] = M.gp * (c0 * v[KBY] - c1 * v[KBX]);
}

template <int SYS, bool ZZ>
__device__ int res_prim2cons_fwd(const Eos& E, const Metric& M, double* v, double* u);

template <int SYS, bool ZZ>
__device__ __forceinline__ int prim2cons(const Eos& E, const Metric& M, double* v, double* u) {
  if constexpr (SysT<SYS>::RES) return res_prim2cons_fwd<SYS, ZZ>(E, M, v, u);
  constexpr bool VISC = SysT<SYS>::VISC;
  constexpr int NV = SysT<SYS>::NV;
  double rh = v[KRH], vx = v[KVX], vy = v[KVY], vz = v[KVZ], pr = v[KPR], en;
  int err = eos_energy(E, rh, pr, en);
  if (err > 0) return err;
  double v2 = M.g1 * vx * vx + M.g2 * vy * vy + M.g3 * vz * vz;
  double glf = 1. / sqrt(1. - v2);
  double w = en + pr;
  double ww = w * (glf * glf);
  u[KRH] = rh * glf;
  if constexpr (VISC) {
    w = w + v[KPIBU];
    ww = w * (glf * glf);
    PiDer d = derived_pi<ZZ>(v, M);
    u[KVX] = (ww * vx + d.pitx);
    u[KVY] = (ww * vy + d.pity);
    u[KVZ] = (ww * vz + d.pitz) * M.g3;
    u[KPR] = ww - pr - v[KPIBU] - d.pitt * M.g0;
    u[KPIBU] = v[KPIBU] * u[KRH];
    u[KPIZZ] = v[KPIZZ] * u[KRH];
    u[KPIXZ] = v[KPIXZ] * u[KRH];
    u[KPIXY] = v[KPIXY] * u[KRH];
    u[KPIYZ] = v[KPIYZ] * u[KRH];
    u[KPIXX] = v[KPIXX] * u[KRH];
    u[KPIYY] = v[KPIYY] * u[KRH];
  } else if constexpr (SysT<SYS>::MHD) {
    double Ef[3], X[3];
    ideal_E(M, v, Ef);
    double E2 = Ef[0] * Ef[0] + Ef[1] * Ef[1] + Ef[2] * Ef[2] / M.g3;
    e_cross_b(M, Ef, v, X);
    double B2 = v[KBX] * v[KBX] + v[KBY] * v[KBY] + v[KBZ] * v[KBZ] * M.g3;
    double EEM = 0.5 * (E2 + B2);
    u[KVX] = ww * vx + X[0];
    u[KVY] = ww * vy + X[1];
    u[KVZ] = ww * vz * M.g3 + X[2];
    u[KPR] = ww - pr + EEM;
    u[KBX] = v[KBX]; u[KBY] = v[KBY]; u[KBZ] = v[KBZ];
    if constexpr (SysT<SYS>::GLM) u[KGLM] = v[KGLM] / M.gp;
  } else {
    u[KVX] = ww * vx;
    u[KVY] = ww * vy;
    u[KVZ] = ww * vz * M.g3;
    u[KPR] = ww - pr;
  }
#pragma unroll
  for (int k = 0; k < NV; k++) u[k] = u[k] * M.gp;
  return 0;
}

// ---------------------------------------------------------------------------
// system_flux split for the HLL solver: side_state evaluates the scalar part
// of system_flux (system.f90:1336-1400 and the fast speeds of 1557-1580) for
// one reconstructed state; u_comp / f_comp then give conserved component k
// and its flux along DIR with the same operations as sys_flux, before the
// final sqrt(-g) factor.  Keeping only these scalars live (instead of four
// nv-vectors) is what keeps the flux kernel's register use low.
// ---------------------------------------------------------------------------
struct SideState {
  double rh, glf, ww, pr, v2, cs2, ca2, vdir, wwdir, enpr;
  PiDer d;
  double E[3], X[3], EEM;   // MHD: covariant E, E x B, field energy
  double Ev[3];             // resistive MHD: contravariant E
};

// resistive systems (defined at the end of this file)
template <int SYS, bool ZZ>
__device__ int res_prim2cons(const Eos& E, const Metric& M, double* v, double* u);
template <int DIR, int SYS, bool ZZ>
__device__ int res_side_state(const Eos& E, const Metric& M, double* v, SideState& s);
template <int SYS>
__device__ double res_u_comp(int k, const Metric& M, const SideState& s, const double* v);
template <int DIR, int SYS>
__device__ double res_f_comp(int k, const Metric& M, const SideState& s, const double* v);
template <int SYS, bool ZZ>
__device__ int res_cons2prim(const Eos& E, const Metric& M, const double* u, double* v, int algo);

template <int DIR, int SYS, bool ZZ>
__device__ __forceinline__ int side_state(const Eos& E, const Metric& M, double* v, SideState& s) {
  if constexpr (SysT<SYS>::RES) return res_side_state<DIR, SYS, ZZ>(E, M, v, s);
  constexpr bool VISC = SysT<SYS>::VISC;
  double rh = v[KRH], vx = v[KVX], vy = v[KVY], vz = v[KVZ], pr = v[KPR], en, cs2;
  int err_sound;
  int err = eos_energy_sound(E, rh, pr, en, cs2, err_sound);
  if (err > 0) return err;
  s.enpr = en + pr;
  double v2 = M.g1 * vx * vx + M.g2 * vy * vy + M.g3 * vz * vz;
  if (v2 > 1) return 333;
  double glf = rsqrt_nr(1. - v2);
  double w = VISC ? en + pr + v[KPIBU] : en + pr;
  s.ww = w * (glf * glf);
  s.rh = rh; s.glf = glf; s.pr = pr; s.v2 = v2;
  s.d = PiDer{0., 0., 0., 0.};
  if constexpr (VISC) s.d = derived_pi<ZZ>(v, M);
  s.ca2 = 0.;
  if constexpr (SysT<SYS>::MHD) {
    ideal_E(M, v, s.E);
    double E2 = s.E[0] * s.E[0] + s.E[1] * s.E[1] + s.E[2] * s.E[2] / M.g3;
    e_cross_b(M, s.E, v, s.X);
    double B2 = v[KBX] * v[KBX] + v[KBY] * v[KBY] + v[KBZ] * v[KBZ] * M.g3;
    s.EEM = 0.5 * (E2 + B2);
    double bb2 = B2 - E2;
    s.ca2 = bb2 / (bb2 + (en + pr));
  }
  s.vdir = DIR == 1 ? vx : (DIR == 2 ? vy : vz);
  s.wwdir = s.ww * s.vdir;
  s.cs2 = cs2;
  return err_sound;
}

// fast speeds lambda+ (lp) and lambda- (lm), system.f90:1557-1580
template <int DIR>
__device__ __forceinline__ void side_speeds(const Metric& M, const SideState& s, double& lp, double& lm) {
  const double gdir = DIR == 1 ? M.g1 : (DIR == 2 ? M.g2 : M.g3);
  const double ca2 = s.ca2;
  double a2 = s.cs2 + ca2 - s.cs2 * ca2;
  double den = 1. / (1. - s.v2 * a2);
  double vf1 = den * (1. - a2) * s.vdir;
  double vf2 = den * sqrt(a2 * (1. - s.v2) * ((1. - s.v2 * a2) / gdir - (1. - a2) * (s.vdir * s.vdir)));
  lp = vf1 + vf2;
  lm = vf1 - vf2;
}

template <int SYS>
__device__ __forceinline__ double u_comp(int k, const Metric& M, const SideState& s, const double* v) {
  if constexpr (SysT<SYS>::RES) return res_u_comp<SYS>(k, M, s, v);
  constexpr bool VISC = SysT<SYS>::VISC;
  if constexpr (SysT<SYS>::MHD) {
    switch (k) {
      case KRH: return s.rh * s.glf;
      case KVX: return s.ww * v[KVX] + s.X[0];
      case KVY: return s.ww * v[KVY] + s.X[1];
      case KVZ: return s.ww * v[KVZ] * M.g3 + s.X[2];
      case KPR: return s.ww - s.pr + s.EEM;
      case KGLM: return v[KGLM] / M.gp;
      default: return v[k];   // B
    }
  }
  switch (k) {
    case KRH: return s.rh * s.glf;
    case KVX: return VISC ? s.ww * v[KVX] + s.d.pitx : s.ww * v[KVX];
    case KVY: return VISC ? s.ww * v[KVY] + s.d.pity : s.ww * v[KVY];
    case KVZ: return VISC ? (s.ww * v[KVZ] + s.d.pitz) * M.g3 : s.ww * v[KVZ] * M.g3;
    case KPR: return VISC ? s.ww - s.pr - v[KPIBU] - s.d.pitt * M.g0 : s.ww - s.pr;
    default: return v[k] * (s.rh * s.glf);
  }
}

template <int DIR, int SYS>
__device__ __forceinline__ double f_comp(int k, const Metric& M, const SideState& s, const double* v) {
  if constexpr (SysT<SYS>::RES) return res_f_comp<DIR, SYS>(k, M, s, v);
  constexpr bool VISC = SysT<SYS>::VISC;
  constexpr bool GLM = SysT<SYS>::GLM;
  const double wd = s.wwdir;
  if (k == KRH) return s.rh * s.glf * s.vdir;
  if constexpr (SysT<SYS>::MHD) {
    const double* E = s.E;
    const double Bx = v[KBX], By = v[KBY], Bz = v[KBZ];
    if constexpr (DIR == 1) {
      switch (k) {
        case KVX: return wd * v[KVX] - E[0] * E[0] - Bx * Bx + s.pr + s.EEM;
        case KVY: return wd * v[KVY] - E[0] * E[1] - Bx * By;
        case KVZ: return wd * v[KVZ] * M.g3 - E[0] * E[2] - Bx * Bz * M.g3;
        case KPR: return wd + s.X[0];
        case KBX: return GLM ? v[KGLM] : 0.;
        case KBY: return -E[2] * M.gm;
        case KBZ: return +E[1] * M.gm;
        default: return v[KBX] / M.gp;
      }
    } else if constexpr (DIR == 2) {
      switch (k) {
        case KVX: return wd * v[KVX] - E[1] * E[0] - By * Bx;
        case KVY: return wd * v[KVY] - E[1] * E[1] - By * By + s.pr + s.EEM;
        case KVZ: return wd * v[KVZ] * M.g3 - E[1] * E[2] - By * Bz * M.g3;
        case KPR: return wd + s.X[1];
        case KBX: return E[2] * M.gm;
        case KBY: return GLM ? v[KGLM] : 0.;
        case KBZ: return -E[0] * M.gm;
        default: return v[KBY] / M.gp;
      }
    } else {
      switch (k) {
        case KVX: return wd * v[KVX] - E[2] / M.g3 * E[0] - Bz * Bx;
        case KVY: return wd * v[KVY] - E[2] / M.g3 * E[1] - Bz * By;
        case KVZ: return wd * v[KVZ] * M.g3 - E[2] * E[2] / M.g3 - Bz * Bz * M.g3 + s.pr + s.EEM;
        case KPR: return wd + s.X[2] / M.g3;
        case KBX: return -E[1] * M.gm;
        case KBY: return E[0] * M.gm;
        case KBZ: return GLM ? v[KGLM] / M.g3 : 0.;
        default: return v[KBZ] / M.gp;
      }
    }
  }
  if (k >= 5) return (v[k] * (s.rh * s.glf)) * s.vdir;
  if constexpr (DIR == 1) {
    switch (k) {
      case KVX: return VISC ? wd * v[KVX] * M.g1 + s.pr + v[KPIBU] + v[KPIXX] * M.g1 : wd * v[KVX] * M.g1 + s.pr;
      case KVY: return VISC ? wd * v[KVY] * M.g2 + v[KPIXY] * M.g2 : wd * v[KVY] * M.g2;
      case KVZ: return VISC ? wd * v[KVZ] * M.g3 + v[KPIXZ] * M.g3 : wd * v[KVZ] * M.g3;
      default: return VISC ? wd + s.d.pitx : wd;
    }
  } else if constexpr (DIR == 2) {
    switch (k) {
      case KVX: return VISC ? wd * v[KVX] * M.g1 + v[KPIXY] * M.g1 : wd * v[KVX] * M.g1;
      case KVY: return VISC ? wd * v[KVY] * M.g2 + s.pr + v[KPIBU] + v[KPIYY] * M.g2 : wd * v[KVY] * M.g2 + s.pr;
      case KVZ: return VISC ? wd * v[KVZ] * M.g3 + v[KPIYZ] * M.g3 : wd * v[KVZ] * M.g3;
      default: return VISC ? wd + s.d.pity : wd;
    }
  } else {
    switch (k) {
      case KVX: return VISC ? wd * v[KVX] * M.g1 + v[KPIXZ] * M.g1 : wd * v[KVX] * M.g1;
      case KVY: return VISC ? wd * v[KVY] * M.g2 + v[KPIYZ] * M.g2 : wd * v[KVY] * M.g2;
      case KVZ: return VISC ? wd * v[KVZ] * M.g3 + s.pr + v[KPIBU] + v[KPIZZ] * M.g3 : wd * v[KVZ] * M.g3 + s.pr;
      default: return VISC ? wd + s.d.pitz : wd;
    }
  }
}

// ---------------------------------------------------------------------------
// cons2prim, system.f90:511-665 (ideal Newton on p; viscous nested
// iteration).  v holds the guess on entry (v[KPR], velocities for viscous).
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Ideal-MHD con2prim: algorithms 4 (ideal gas), 5 (p = e/3) and 6 (quartic,
// e = 3p, Minkowski norms) of system_cons2prim (system.f90:880-1158) with
// rtsafe (2242-2290) and the cubic solvers solver7 / solver8 (2102-2201).
// ---------------------------------------------------------------------------
struct MhdPar { double B2, SB2, S2, UU, D; };

template <bool IDEALGAS>
__device__ __forceinline__ void mhd_cubic(const MhdPar& P, double x, double& f, double& df, double& wout) {
  const double third = 1.0 / 3.0;
  const double gam = IDEALGAS ? (GAMMA_AD - 1) / GAMMA_AD : 0.25;
  double v2 = x;
  double rho = IDEALGAS ? P.D * sqrt(1.0 - v2) : 0.;
  double c3 = 1.0 - gam * (1.0 - v2);
  double c2 = IDEALGAS ? gam * rho + .5 * P.B2 * (1.0 + v2) - P.UU : .5 * P.B2 * (1.0 + v2) - P.UU;
  double c0 = -0.5 * P.SB2;
  double w;
  if (!IDEALGAS && c0 == 0.0) {
    w = -c2 / c3;
  } else {
    w = dmax(-c2 / c3, pow(-c0 / c3, third));
    const int nit = IDEALGAS ? 50 : 100;
    const double tol = IDEALGAS ? 1.e-10 : 1.e-12;
    for (int iter = 1; iter <= nit; iter++) {
      double dw = -((c3 * w + c2) * (w * w) + c0) / ((3 * c3 * w + 2 * c2) * w);
      if (fabs(dw / w) < tol) break;
      w = w + dw;
    }
  }
  double dc3 = gam;
  double dc2 = IDEALGAS ? .5 * (P.B2 - gam * rho / (1. - v2)) : .5 * P.B2;
  double dlogw = -(dc3 * w + dc2) / (3 * c3 * w + 2 * c2);
  double wb = w + P.B2;
  double vb2 = P.SB2 / (w * w);
  f = (wb * wb) * v2 - (2 * w + P.B2) * vb2 - P.S2;
  df = wb * (wb + 2 * dlogw * (w * v2 + vb2));
  wout = w;
}

template <bool IDEALGAS>
__device__ __forceinline__ double mhd_rtsafe(const MhdPar& P, double x0, double x1, double x2,
                                             double xacc, int& ierr, double& w) {
  double f, df, xl = x1, xh = x2, r = x0, dxold = fabs(x2 - x1), dx = dxold, temp;
  ierr = 0;
  mhd_cubic<IDEALGAS>(P, r, f, df, w);
  for (int j = 1; j <= 100; j++) {
    if (((r - xh) * df - f) * ((r - xl) * df - f) > 0. || fabs(2. * f) > fabs(dxold * df)) {
      dxold = dx;
      dx = 0.5 * (xh - xl);
      r = xl + dx;
      if (xl == r) { mhd_cubic<IDEALGAS>(P, r, f, df, w); return r; }
    } else {
      dxold = dx;
      dx = f / df;
      temp = r;
      r = r - dx;
      if (temp == r) { mhd_cubic<IDEALGAS>(P, r, f, df, w); return r; }
    }
    if (fabs(dx) < xacc) return r;
    mhd_cubic<IDEALGAS>(P, r, f, df, w);
    if (f < 0.) xl = r; else xh = r;
  }
  ierr = 1;
  return r;
}

__device__ __forceinline__ double get_x9(double y, double s2, double b2, double sb2) {
  return (s2 * (y * y) + b2 * sb2 + 2 * sb2 * y) / ((y * y) * ((b2 * b2) + 2 * b2 * y + (y * y)));
}

// Brent's method for N nonlinear equations (ACM algorithm 554, the
// reference's vendored BRENTM, libacm554.f90:8-359), with the Jacobian
// estimated row by row from divided differences and reduced by Householder
// steps.  fcn(x, f, k) evaluates component k.  N <= 3, so every array lives
// in registers (the loops over components are unrolled).  Returns INFO.
template <int N, typename Fcn>
__device__ __forceinline__ int brentm(Fcn&& fcn, double* x, double ftol, double xtol, int maxfev, int mopt) {
  const double p05 = 5.e-2, scale = 1.e1;
  const double eps = 1.4901161193847656e-08;   // sqrt(EPSILON(1.d0))
  double q[N][N], sigma[N], wa1[N], wa2[N], fvec[N];
  int info = 0, nfcall = 0, nfev = 0;
  int nier6 = -1, nier7 = -1, nier8 = 0;
  double fnorm = 0., difit = 0., xnorm = 0.;
#pragma unroll
  for (int i = 0; i < N; i++) xnorm = dmax(xnorm, fabs(x[i]));
  double delta = scale * xnorm;
  if (xnorm == 0.) delta = scale;
  for (;;) {
    int nsing = N;
    double fnorm1 = fnorm, difit1 = difit;
    fnorm = 0.;
    double h = eps * xnorm;
    if (h == 0.) h = eps;
#pragma unroll
    for (int j = 0; j < N; j++) {
#pragma unroll
      for (int i = 0; i < N; i++) q[j][i] = 0.;
      q[j][j] = h;
      wa1[j] = x[j];
    }
#pragma unroll
    for (int k = 0; k < N; k++) {
      fcn(wa1, fvec, k);
      const double fky = fvec[k];
      nfcall++;
#pragma unroll
      for (int j = k; j < N; j++) {
#pragma unroll
        for (int i = 0; i < N; i++) wa2[i] = wa1[i] + q[j][i];
        fcn(wa2, fvec, k);
        nfcall++;
        sigma[j] = fvec[k] - fky;
      }
      nfev = nfcall / N;
      fnorm = dmax(fnorm, fabs(fky));
      fvec[k] = fky;
      double eta = 0.;
#pragma unroll
      for (int i = k; i < N; i++) eta = dmax(eta, fabs(sigma[i]));
      if (eta == 0.) continue;
      nsing = nsing - 1;
      double sknorm = 0.;
#pragma unroll
      for (int i = k; i < N; i++) {
        sigma[i] = sigma[i] / eta;
        sknorm = sknorm + sigma[i] * sigma[i];
      }
      sknorm = sqrt(sknorm);
      if (sigma[k] < 0.) sknorm = -sknorm;
      sigma[k] = sigma[k] + sknorm;
#pragma unroll
      for (int i = 0; i < N; i++) wa2[i] = 0.;
#pragma unroll
      for (int j = k; j < N; j++) {
#pragma unroll
        for (int i = 0; i < N; i++) wa2[i] = wa2[i] + sigma[j] * q[j][i];
      }
#pragma unroll
      for (int j = k; j < N; j++) {
        const double temp = sigma[j] / (sknorm * sigma[k]);
#pragma unroll
        for (int i = 0; i < N; i++) q[j][i] = q[j][i] - temp * wa2[i];
      }
      sigma[k] = sknorm * eta;
      double temp = fky / sigma[k];
      if (h * fabs(temp) > delta) temp = copysign(delta / h, temp);
#pragma unroll
      for (int i = 0; i < N; i++) wa1[i] = wa1[i] + temp * q[k][i];
    }
    xnorm = 0.;
    difit = 0.;
#pragma unroll
    for (int i = 0; i < N; i++) {
      xnorm = dmax(xnorm, fabs(wa1[i]));
      difit = dmax(difit, fabs(x[i] - wa1[i]));
      x[i] = wa1[i];
    }
    delta = dmax(delta, scale * xnorm);
    const bool conv = (fnorm < fnorm1 && difit < difit1 && nsing == 0);
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
    if (nsing == N) info = 5;
    if (nier6 == 5) info = 6;
    if (nier7 == 3) info = 7;
    if (nier8 == 4) info = 8;
    if (info != 0) return info;
    if (!conv || difit > p05 * xnorm || mopt == 1) continue;
    for (int m = 2; m <= mopt; m++) {       // iterative refinement
      fnorm1 = fnorm;
      fnorm = 0.;
      bool stop = false;
#pragma unroll
      for (int k = 0; k < N; k++) {
        if (stop) break;
        fcn(wa1, fvec, k);
        const double fky = fvec[k];
        nfcall++;
        nfev = nfcall / N;
        fnorm = dmax(fnorm, fabs(fky));
        if (!(fnorm < fnorm1)) {
          fnorm = fnorm1;
          stop = true;
          break;
        }
        const double temp = fky / sigma[k];
#pragma unroll
        for (int i = 0; i < N; i++) wa1[i] = wa1[i] + temp * q[k][i];
      }
      if (stop) break;
      xnorm = 0.;
      difit = 0.;
#pragma unroll
      for (int i = 0; i < N; i++) {
        xnorm = dmax(xnorm, fabs(wa1[i]));
        difit = dmax(difit, fabs(x[i] - wa1[i]));
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

template <int SYS>
__device__ __forceinline__ int mhd_cons2prim(const Eos& E, int algo, const Metric& M, const double* u, double* v) {
  constexpr int NV = SysT<SYS>::NV;
  const double gm = M.gm, g3 = M.g3;
  double uin[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) uin[k] = gm * u[k];
  const double D = uin[KRH], sx = uin[KVX], sy = uin[KVY], sz = uin[KVZ], UU = uin[KPR];
  const double bx = uin[KBX], by = uin[KBY], bz = uin[KBZ];
  const double SB = (sx * bx + sy * by + sz * bz);
  const double SB2 = SB * SB;
  if (algo == 1 || algo == 2) {
    // system.f90:668-793: BRENTM on the 3x3 system sist3 (v^2, W, p) or the
    // 2x2 system sist4 (W, p; v^2 explicit), system.f90:1199-1266, warm
    // started from the cell's previous primitives.  The error code of the
    // lookup of the previous energy density is the routine's error code.
    const double B2 = bx * bx + by * by + bz * bz * g3;
    const double SS2 = sx * sx + sy * sy + sz * sz / g3;
    double pr = v[KPR], en = 0.;
    const int err0 = eos_energy(E, v[KRH], pr, en);
    const double v2 = M.g1 * (v[KVX] * v[KVX]) + M.g2 * (v[KVY] * v[KVY]) + g3 * (v[KVZ] * v[KVZ]);
    double xs, ys;
    if (algo == 1) {
      double X[3] = {v2, (en + pr) / (1 - v2), pr};
      brentm<3>([&](const double* Xk, double* F, int k) {
        const double xx = Xk[0], yy = Xk[1], pp = Xk[2];
        if (k == 0) F[0] = (yy + B2) * (yy + B2) * xx - SB2 / (yy * yy) * (2 * yy + B2) - SS2;
        if (k == 1) F[1] = yy - pp + 0.5 * (1. + xx) * B2 - 0.5 * SB2 / (yy * yy) - UU;
        if (k == 2) {
          double e3 = 0.;
          eos_energy(E, D * sqrt(1 - xx), pp, e3);
          F[2] = yy - (pp + e3) / (1 - xx);
        }
      }, X, 1.e-14, 1.e-14, 300, 3);
      xs = X[0]; ys = X[1]; pr = X[2];
    } else {
      double X[2] = {(en + pr) / (1 - v2), pr};
      double xin = 0.;
      brentm<2>([&](const double* Xk, double* F, int k) {
        const double yy = Xk[0], pp = Xk[1];
        const double xx = (SS2 + SB2 * (2. * yy + B2) / (yy * yy)) / ((yy + B2) * (yy + B2));
        xin = xx;
        if (k == 1) {
          double e3 = 0.;
          eos_energy(E, D * sqrt(1 - xx), pp, e3);
          F[1] = yy - (pp + e3) / (1 - xx);
        }
        if (k == 0) F[0] = yy - pp + 0.5 * (1. + xx) * B2 - 0.5 * SB2 / (yy * yy) - UU;
      }, X, 1.e-14, 1.e-14, 300, 3);
      ys = X[0]; pr = X[1]; xs = xin;   // v^2 of the last sist4 call
    }
    v[KBX] = bx; v[KBY] = by; v[KBZ] = bz;
    v[KRH] = D * sqrt(1 - xs);
    v[KVX] = (sx + SB * bx / ys) / (ys + B2);
    v[KVY] = (sy + SB * by / ys) / (ys + B2);
    v[KVZ] = (sz / g3 + SB * bz / ys) / (ys + B2);
    v[KPR] = pr;
    if constexpr (SysT<SYS>::GLM) v[KGLM] = u[KGLM];
    return err0;
  }
  if (algo == 4 || algo == 5) {
    MhdPar P;
    P.S2 = sx * sx + sy * sy + sz * sz / g3;
    P.B2 = bx * bx + by * by + bz * bz * g3;
    P.SB2 = SB2; P.UU = UU; P.D = D;
    double v2, w;
    int ierr;
    if (algo == 4) {
      double v2_old = v[KVX] * v[KVX] + v[KVY] * v[KVY] + v[KVZ] * v[KVZ];
      v2 = mhd_rtsafe<true>(P, v2_old, 0.0, 1.0 - 1.e-11, 1.e-12, ierr, w);
      if (ierr > 0) return 0;   // the reference returns with errcode 0, v untouched
#pragma unroll
      for (int k = 0; k < NV; k++) v[k] = uin[k];
      if constexpr (SysT<SYS>::GLM) v[KGLM] = u[KGLM];
      double rh = D * sqrt(1. - v2);
      double pg = (GAMMA_AD - 1) / GAMMA_AD * (w * (1. - v2) - rh);
      v[KRH] = rh;
      v[KPR] = dmax(pg, 1.e-12);
    } else {
      double v2_old = v[KVX] * v[KVX] + v[KVY] * v[KVY] + v[KVZ] * v[KVZ] * g3;
      v2 = mhd_rtsafe<false>(P, v2_old, 0.0, 1.0 - 1.e-11, 1.e-12, ierr, w);
      if ((v2 >= 1) || (ierr > 0)) return 778;
#pragma unroll
      for (int k = 0; k < NV; k++) v[k] = uin[k];
      if constexpr (SysT<SYS>::GLM) v[KGLM] = u[KGLM];
      v[KPR] = 0.25 * (1 - v2) * w;
      v[KRH] = D * sqrt(1. - v2);
    }
    double vb = SB / w;
    v[KVX] = (sx + vb * bx) / (w + P.B2);
    v[KVY] = (sy + vb * by) / (w + P.B2);
    v[KVZ] = (sz / g3 + vb * bz) / (w + P.B2);
    return 0;
  }
  if (algo == 6) {
    // exact quartic for e = 3p (system.f90:1000-1158)
    double v2_old = v[KVX] * v[KVX] + v[KVY] * v[KVY] + v[KVZ] * v[KVZ];
    double oldsol9 = 4 * v[KPR] / (1 - v2_old);
    double S2 = sx * sx + sy * sy + sz * sz;
    double B2 = bx * bx + by * by + bz * bz;
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
    bool exact9 = true;
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
      // solver10_approximate, system.f90:2205-2223
      double xx = oldsol9;
      int iter;
      for (iter = 1; iter <= 100; iter++) {
        double dx = -(12 * (xx * xx * xx) + 3 * b9 * (xx * xx) + 2 * c9 * xx + d9);
        if (fabs(dx / xx) < 1.e-12) break;
        xx = xx + dx;
      }
      if (iter == 100) return 12;
      sol9 = xx;
      exact9 = false;
      v2 = get_x9(sol9, S2, B2, SB2);
      if ((v2 < 0) || (v2 > 1) || (sol9 < 0)) return 12;
    }
    if (exact9) {
      double rad129 = -4 * (sbig9 * sbig9) - 2 * p9 + q9 / sbig9;
      double rad349 = -4 * (sbig9 * sbig9) - 2 * p9 - q9 / sbig9;
      double rapp9 = -b9 / 12;
      double sol1 = -1, sol2 = -1, sol3 = -1, sol4 = -1, dist9 = 1.e100, v9;
      if (rad129 >= 0) { double r = 0.5 * sqrt(rad129); sol1 = rapp9 - sbig9 - r; sol2 = rapp9 - sbig9 + r; }
      if (rad349 >= 0) { double r = 0.5 * sqrt(rad349); sol3 = rapp9 + sbig9 - r; sol4 = rapp9 + sbig9 + r; }
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
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = uin[k];
    v[KRH] = D * sqrt(1. - v2);
    v[KPR] = sol9 / 4 * (1 - v2);
    double vb = SB / sol9;
    v[KVX] = (sx + vb * bx) / (sol9 + B2);
    v[KVY] = (sy + vb * by) / (sol9 + B2);
    v[KVZ] = (sz + vb * bz) / (sol9 + B2);
    if constexpr (SysT<SYS>::GLM) v[KGLM] = u[KGLM];
    return 0;
  }
  return 11;
}

template <int SYS, bool ZZ>
__device__ __forceinline__ int cons2prim(const Eos& E, const Metric& M, const double* u, double* v,
                                         int algo = 5) {
  if constexpr (SysT<SYS>::RES) return res_cons2prim<SYS, ZZ>(E, M, u, v, algo);
  constexpr bool VISC = SysT<SYS>::VISC;
  constexpr int NV = SysT<SYS>::NV;
  if constexpr (SysT<SYS>::MHD) return mhd_cons2prim<SYS>(E, algo, M, u, v);
  const double tol = 1.e-12, eps = 1.e-12;
  const double gm = M.gm;
  int errcode = 0;
  double d = gm * u[KRH];
  double sx = gm * u[KVX];
  double sy = gm * u[KVY];
  double sz = gm * u[KVZ] / M.g3;
  double et = gm * u[KPR];
  double s2 = M.g1 * sx * sx + M.g2 * sy * sy + M.g3 * sz * sz;
  // eos_pressure leaves its outputs unset on an error (out of table range)
  // and the Newton step still uses them (system.f90:534-538, 627-631);
  // before the first successful call the reference reads uninitialised
  // locals there.  They start at 0 here.
  double pr, w = 0., v2, glf, rh = 0., en, drhdpr, dprdrh = 0., dprden = 0., f, df, dpr, pr_eos = 0.;
  if constexpr (!VISC) {
    pr = v[KPR];
    for (int iter = 1; iter <= 100; iter++) {
      w = et + pr;
      v2 = s2 / (w * w);
      if (v2 >= 1.) return 777;
      glf = 1. / sqrt(1. - v2);
      rh = d / glf;
      en = w * (1. - v2) - pr;
      drhdpr = rh * v2 * glf / (en + pr);
      errcode = eos_pressure(E, rh, en, pr_eos, dprdrh, dprden);
      f = pr_eos - pr;
      df = dprdrh * drhdpr + dprden * v2 - 1.;
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
  } else {
    double uin[NV], vprim[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) uin[k] = gm * u[k];
    d = uin[KRH];
#pragma unroll
    for (int k = KPIBU; k <= KPIZZ; k++) v[k] = uin[k] / uin[KRH];
    double vx = v[KVX], vy = v[KVY], vz = v[KVZ];
#pragma unroll
    for (int k = 0; k < NV; k++) vprim[k] = v[k];
    for (int iv = 1; iv <= 100; iv++) {
      vprim[KVX] = vx; vprim[KVY] = vy; vprim[KVZ] = vz;
      PiDer q = derived_pi<ZZ>(vprim, M);
      double en1 = et + v[KPIBU] + q.pitt * M.g0;
      double sx_cov = uin[KVX] - q.pitx;
      double sy_cov = uin[KVY] - q.pity;
      double sz_cov = uin[KVZ] - q.pitz * M.g3;
      sx = sx_cov;
      sy = sy_cov;
      sz = sz_cov / M.g3;
      s2 = sx * sx_cov + sy * sy_cov + sz * sz_cov;
      pr = v[KPR];
      for (int ip = 1; ip <= 100; ip++) {
        double qq = en1 + pr;
        v2 = s2 / (qq * qq);
        if (v2 >= 1.) return 107;
        glf = 1. / sqrt(1. - v2);
        rh = d / glf;
        en = (en1 + pr) * (1. - v2) - pr - vprim[KPIBU];
        drhdpr = rh * v2 * glf / (en + pr);
        errcode = eos_pressure(E, rh, en, pr_eos, dprdrh, dprden);
        f = pr_eos - pr;
        df = dprdrh * drhdpr + dprden * v2 - 1.;
        dpr = -f / df;
        if (fabs(dpr) < tol * pr) break;
        pr = pr + dpr;
      }
      vx = sx / (en1 + pr); if (fabs(vx) < eps) vx = 0.;
      vy = sy / (en1 + pr); if (fabs(vy) < eps) vy = 0.;
      vz = sz / (en1 + pr); if (fabs(vz) < eps) vz = 0.;
      vprim[KVX] = vx; vprim[KVY] = vy; vprim[KVZ] = vz;
      if (fabs(M.g1 * (vx * vx) + M.g2 * (vy * vy) + M.g3 * (vz * vz) - v2) < eps) break;
    }
    v[KRH] = rh; v[KVX] = vx; v[KVY] = vy; v[KVZ] = vz; v[KPR] = pr;
    if constexpr (ZZ) derived_pi<true>(v, M);
    return errcode;
  }
}

// ---------------------------------------------------------------------------
// Resistive MHD (E evolved, Dedner scalar), without and with viscosity.  The
// reference has no working path for it (system_cons2prim has no algorithm 7,
// system.f90:1160-1162; viscosity with MHD is refused, common.f90:1904-1908);
// this is the build's design, assembled from the reference's pieces: the
// fluid part of every operator is the hydro or viscous system's own, and the
// electromagnetic part follows the MHD branch of system_prim2cons /
// system_flux (system.f90:200-263, 1425-1549) with E evolved: E x B and
// (E^2 + B^2)/2 in the momentum and energy, the Maxwell stresses in the
// momentum fluxes, the induction (B) and Ampere (E) fluxes, and the GLM
// fluxes of psi and the normal B.  Covariant E_eta = g33 E^eta throughout
// (the reference's resistive flux drops g33, system.f90:1369).
// ---------------------------------------------------------------------------
struct EmTerms {
  double Ec[3];   // covariant E
  double X[3];    // E x B, covariant, times sqrt(-g) (system_EXB)
  double E2, B2, eem;
};

__device__ __forceinline__ EmTerms em_terms(const Metric& M, const double* Ev, const double* B) {
  EmTerms t;
  t.Ec[0] = Ev[0];
  t.Ec[1] = Ev[1];
  t.Ec[2] = Ev[2] * M.g3;
  t.E2 = t.Ec[0] * Ev[0] + t.Ec[1] * Ev[1] + t.Ec[2] * Ev[2];
  t.B2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2] * M.g3;
  t.X[0] = M.gp * (Ev[1] * B[2] - Ev[2] * B[1]);
  t.X[1] = M.gp * (Ev[2] * B[0] - Ev[0] * B[2]);
  t.X[2] = M.gp * (Ev[0] * B[1] - Ev[1] * B[0]);
  t.eem = 0.5 * (t.E2 + t.B2);
  return t;
}

// contravariant a x b of contravariant a, b (minus the E of
// system_compute_E_in_ideal_MHD, system.f90:1882-1897, made contravariant)
__device__ __forceinline__ void cross_contra(const Metric& M, const double* a, const double* b, double* c) {
  c[0] = M.gp * (a[1] * b[2] - a[2] * b[1]);
  c[1] = M.gp * (a[2] * b[0] - a[0] * b[2]);
  c[2] = M.gp * (a[0] * b[1] - a[1] * b[0]) / M.g3;
}

template <int SYS, bool ZZ>
__device__ int res_prim2cons(const Eos& E, const Metric& M, double* v, double* u) {
  using T = SysT<SYS>;
  const int e = prim2cons<T::FLUID, ZZ>(E, M, v, u);
  if (e > 0) return e;
  const EmTerms t = em_terms(M, &v[T::KE], &v[T::KB]);
  u[KVX] = u[KVX] + M.gp * t.X[0];
  u[KVY] = u[KVY] + M.gp * t.X[1];
  u[KVZ] = u[KVZ] + M.gp * t.X[2];
  u[KPR] = u[KPR] + M.gp * t.eem;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    u[T::KB + k] = M.gp * v[T::KB + k];
    u[T::KE + k] = M.gp * v[T::KE + k];
  }
  u[T::KG] = v[T::KG];   // (psi / gp) * gp, system.f90:253
  return 0;
}

template <int SYS, bool ZZ>
__device__ int res_prim2cons_fwd(const Eos& E, const Metric& M, double* v, double* u) {
  return res_prim2cons<SYS, ZZ>(E, M, v, u);
}

// the fluid state's side values, then the field's; the fluid variables take
// the fast magnetosonic speeds of system.f90:1557-1580 with B^2 - E^2 for B^2
template <int DIR, int SYS, bool ZZ>
__device__ int res_side_state(const Eos& E, const Metric& M, double* v, SideState& s) {
  using T = SysT<SYS>;
  const int e = side_state<DIR, T::FLUID, ZZ>(E, M, v, s);
  if (e) return e;
  const EmTerms t = em_terms(M, &v[T::KE], &v[T::KB]);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    s.Ev[k] = v[T::KE + k];
    s.E[k] = t.Ec[k];
    s.X[k] = t.X[k];
  }
  s.EEM = t.eem;
  const double bb2 = dmax(t.B2 - t.E2, 0.);
  s.ca2 = bb2 / (bb2 + s.enpr);
  return 0;
}

template <int SYS>
__device__ double res_u_comp(int k, const Metric& M, const SideState& s, const double* v) {
  using T = SysT<SYS>;
  if (k < T::NF) {
    double r = u_comp<T::FLUID>(k, M, s, v);
    if (k == KVX) r = r + s.X[0];
    else if (k == KVY) r = r + s.X[1];
    else if (k == KVZ) r = r + s.X[2];
    else if (k == KPR) r = r + s.EEM;
    return r;
  }
  if (k == T::KG) return v[T::KG] / M.gp;
  return v[k];   // B, E
}

template <int DIR, int SYS>
__device__ double res_f_comp(int k, const Metric& M, const SideState& s, const double* v) {
  using T = SysT<SYS>;
  const double* B = &v[T::KB];
  const double* Ev = s.Ev;
  const double* Ec = s.E;
  constexpr int D = DIR - 1;
  if (k < T::NF) {
    const double r = f_comp<DIR, T::FLUID>(k, M, s, v);
    // Maxwell stress -E^d E_i - B^d B_i + delta_di EEM and the Poynting flux
    double em = 0.;
    if (k == KVX) em = -Ev[D] * Ec[0] - B[D] * B[0] + (D == 0 ? s.EEM : 0.);
    else if (k == KVY) em = -Ev[D] * Ec[1] - B[D] * B[1] + (D == 1 ? s.EEM : 0.);
    else if (k == KVZ) em = -Ev[D] * Ec[2] - B[D] * B[2] * M.g3 + (D == 2 ? s.EEM : 0.);
    else if (k == KPR) em = D == 2 ? s.X[2] / M.g3 : s.X[D];
    else return r;
    return r + em;
  }
  const double psi = v[T::KG];
  if constexpr (DIR == 1) {
    if (k == T::KB) return psi;
    if (k == T::KB + 1) return -Ec[2] * M.gm;
    if (k == T::KB + 2) return Ec[1] * M.gm;
    if (k == T::KE) return 0.;
    if (k == T::KE + 1) return B[2] * M.g3 * M.gm;
    if (k == T::KE + 2) return -B[1] * M.g2 * M.gm;
    return B[0] / M.gp;
  } else if constexpr (DIR == 2) {
    if (k == T::KB) return Ec[2] * M.gm;
    if (k == T::KB + 1) return psi;
    if (k == T::KB + 2) return -Ec[0] * M.gm;
    if (k == T::KE) return -B[2] * M.g3 * M.gm;
    if (k == T::KE + 1) return 0.;
    if (k == T::KE + 2) return B[0] * M.g1 * M.gm;
    return B[1] / M.gp;
  } else {
    if (k == T::KB) return -Ec[1] * M.gm;
    if (k == T::KB + 1) return Ec[0] * M.gm;
    if (k == T::KB + 2) return psi / M.g3;
    if (k == T::KE) return B[1] * M.g2 * M.gm;
    if (k == T::KE + 1) return -B[0] * M.g1 * M.gm;
    if (k == T::KE + 2) return 0.;
    return B[2] / M.gp;
  }
}

// E and B are read off the conserved variables, their energy and momentum
// subtracted, and the fluid system's inversion runs on the rest
template <int SYS, bool ZZ>
__device__ int res_cons2prim(const Eos& E, const Metric& M, const double* u, double* v, int algo) {
  using T = SysT<SYS>;
  double B[3], Ev[3], uf[T::NF];
#pragma unroll
  for (int k = 0; k < 3; k++) { B[k] = u[T::KB + k] * M.gm; Ev[k] = u[T::KE + k] * M.gm; }
  const EmTerms t = em_terms(M, Ev, B);
#pragma unroll
  for (int k = 0; k < T::NF; k++) uf[k] = u[k];
  uf[KVX] = u[KVX] - M.gp * t.X[0];
  uf[KVY] = u[KVY] - M.gp * t.X[1];
  uf[KVZ] = u[KVZ] - M.gp * t.X[2];
  uf[KPR] = u[KPR] - M.gp * t.eem;
  const int e = cons2prim<T::FLUID, ZZ>(E, M, uf, v, algo);
#pragma unroll
  for (int k = 0; k < 3; k++) { v[T::KB + k] = B[k]; v[T::KE + k] = Ev[k]; }
  v[T::KG] = u[T::KG];
  return e;
}

// The stiff current of Ohm's law per unit sqrt(-g) (system.f90:1653-1662,
// with v x B in the metric-consistent form and v.B computed):
// -gamma [sig (E - (v.E) v + v x B) + sch (B - (v.B) v - v x E)]
__device__ __forceinline__ void ohm_stiff(const Metric& M, const double* vel, const double* Ev, const double* B,
                                          double sig, double sch, double* S) {
  const double v2 = M.g1 * vel[0] * vel[0] + M.g2 * vel[1] * vel[1] + M.g3 * vel[2] * vel[2];
  const double glf = 1. / sqrt(1. - v2);
  double vxB[3], vxE[3];
  cross_contra(M, vel, B, vxB);
  cross_contra(M, vel, Ev, vxE);
  const double vE = vel[0] * Ev[0] + vel[1] * Ev[1] + M.g3 * vel[2] * Ev[2];
  const double vB = vel[0] * B[0] + vel[1] * B[1] + M.g3 * vel[2] * B[2];
#pragma unroll
  for (int k = 0; k < 3; k++)
    S[k] = -glf * (sig * (Ev[k] - vE * vel[k] + vxB[k]) + sch * (B[k] - vB * vel[k] - vxE[k]));
}

// The implicit part of an IMEX stage: u holds U*; finds E with
// E = E* + adt R(E, v), R the stiff current above and v the fluid velocity
// of the inversion with that E (B and the fluid's D, S, energy do not
// change), by fixed-point iteration on v around a 3x3 linear solve for E
// (Cramer's rule), until E changes by at most 1e-12 of the field scale
// max(|E|, |B|); the fluid inversion of the last E before that is kept (one
// inversion per cell when the first update of E already converges).  Writes
// the new E into u.  The role of the reference's
// missing algorithm 7 (its locals, system.f90:494-497, declare a 3x3 system).
template <int SYS, bool ZZ>
__device__ int res_cons2prim_implicit(const Eos& E, const Metric& M, double* u, double* v, double adt,
                                      double sig, double sch) {
  using T = SysT<SYS>;
  double B[3], Es[3], Ev[3], uf[T::NF];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    B[k] = u[T::KB + k] * M.gm;
    Es[k] = u[T::KE + k] * M.gm;
    Ev[k] = v[T::KE + k];
  }
  // every fluid inversion starts from the pressure and velocity held on
  // entry, as the one inversion of an explicit stage does: the viscous
  // inversion stops its velocity iteration at an absolute tolerance, so
  // chaining the inversions would carry that iteration on and make the
  // result depend on where the loop stops
  const double g0x = v[KVX], g0y = v[KVY], g0z = v[KVZ], g0p = v[KPR];
  int e = 0;
  const double gp = M.gp, g3 = M.g3;
  for (int it = 0; it < 50; it++) {
    const double vel[3] = {v[KVX], v[KVY], v[KVZ]};
    const double v2 = M.g1 * vel[0] * vel[0] + M.g2 * vel[1] * vel[1] + M.g3 * vel[2] * vel[2];
    const double glf = 1. / sqrt(1. - v2);
    const double a = adt * glf * sig, b = adt * glf * sch;
    double vxB[3];
    cross_contra(M, vel, B, vxB);
    const double vB = vel[0] * B[0] + vel[1] * B[1] + g3 * vel[2] * B[2];
    const double vl[3] = {vel[0], vel[1], g3 * vel[2]};
    double r[3];
#pragma unroll
    for (int i = 0; i < 3; i++) r[i] = Es[i] - a * vxB[i] - b * (B[i] - vB * vel[i]);
    // (1 + a) E - a (v.E) v - b (v x E) = r;  (v x E)^i = C_ij E^j
    const double C[3][3] = {{0., -gp * vel[2], gp * vel[1]},
                            {gp * vel[2], 0., -gp * vel[0]},
                            {-gp * vel[1] / g3, gp * vel[0] / g3, 0.}};
    double m[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) m[i][j] = (i == j ? 1. + a : 0.) - a * vel[i] * vl[j] - b * C[i][j];
    const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                       m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                       m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    double En[3];
    En[0] = (r[0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (r[1] * m[2][2] - m[1][2] * r[2]) +
             m[0][2] * (r[1] * m[2][1] - m[1][1] * r[2])) / det;
    En[1] = (m[0][0] * (r[1] * m[2][2] - m[1][2] * r[2]) - r[0] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
             m[0][2] * (m[1][0] * r[2] - r[1] * m[2][0])) / det;
    En[2] = (m[0][0] * (m[1][1] * r[2] - r[1] * m[2][1]) - m[0][1] * (m[1][0] * r[2] - r[1] * m[2][0]) +
             r[0] * (m[1][0] * m[2][1] - m[1][1] * m[2][0])) / det;
    // the change of E against the field scale max(|E|, |B|), not against |E|
    // itself, which can be orders of magnitude smaller than B
    double dE = 0., mE = 0.;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      dE = dmax(dE, fabs(En[k] - Ev[k]));
      mE = dmax(mE, dmax(fabs(En[k]), fabs(B[k])));
      Ev[k] = En[k];
    }
    // converged: v is the inversion with the previous E, equal to E to 1e-12
    // of the field scale
    if (it > 0 && dE <= 1e-12 * mE) break;
    const EmTerms t = em_terms(M, Ev, B);
#pragma unroll
    for (int k = 0; k < T::NF; k++) uf[k] = u[k];
    uf[KVX] = u[KVX] - gp * t.X[0];
    uf[KVY] = u[KVY] - gp * t.X[1];
    uf[KVZ] = u[KVZ] - gp * t.X[2];
    uf[KPR] = u[KPR] - gp * t.eem;
    v[KVX] = g0x; v[KVY] = g0y; v[KVZ] = g0z; v[KPR] = g0p;
    e = cons2prim<T::FLUID, ZZ>(E, M, uf, v, 7);
    if (e) break;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    v[T::KB + k] = B[k];
    v[T::KE + k] = Ev[k];
    u[T::KE + k] = Ev[k] * gp;
  }
  v[T::KG] = u[T::KG];
  return e;
}

}  // namespace echo
