// Source terms of the grid update: Bjorken geometric source (ideal HD) and
// second-order viscous (Israel-Stewart) relaxation, system.f90:1587-1842,
// with the shear tensor of viscous.f90:283-393 and the transport
// coefficients of viscous.f90:172-199.
#pragma once
#include "phys.hpp"

namespace echo {

// The 16 velocity derivatives d_mu u^nu used by system_source, in the order
// dutdt dutdx dutdy dutdz duxdx duydy duzdz duzdt duzdx duzdy duxdt duxdy
// duxdz duydt duydx duydz.
struct Deriv {
  double dutdt, dutdx, dutdy, dutdz, duxdx, duydy, duzdz, duzdt, duzdx, duzdy;
  double duxdt, duxdy, duxdz, duydt, duydx, duydz;
};

struct ViscParams {
  double eta_over_s, tau_pi_coeff;
  int bulkvis, init_type;
  // bulk viscosity zeta/s peaked at T_c (echo_set_bulk_peak), else the
  // reference's zeta = 2 eta (1/3 - c_s^2)
  int zeta_peak;
  double zs_max, zs_tc, zs_w;
};

// bulk viscosity from eta, c_s^2, T (GeV) and s: viscous.f90:186-190, or
// zeta = hbar zs_max exp(-(T - T_c)^2 / (2 w^2)) s with echo_set_bulk_peak
__device__ __forceinline__ double bulk_zeta(const ViscParams& P, double eta, double cs2, double temp, double s) {
  if (!P.bulkvis) return 0.;
  if (P.zeta_peak) {
    const double x = (temp - P.zs_tc) / P.zs_w;
    return 0.197326 * P.zs_max * exp(-0.5 * x * x) * s;   // hbar c, common.f90:47
  }
  return 2. * eta * (1. / 3. - cs2);
}

enum { SXX = 0, SYY = 1, SZZ = 2, SXY = 3, SXZ = 4, SYZ = 5 };

// get_sigma, viscous.f90:283-393 (cell primitives v, derivatives D)
__device__ __forceinline__ void get_sigma(int coordinates, const Metric& M, const double* v,
                                          const Deriv& D, double* sigma) {
  double ux = v[KVX], uy = v[KVY], uz = v[KVZ];
  double ut = 1. / sqrt(1. - M.g1 * ux * ux - M.g2 * uy * uy - M.g3 * uz * uz);
  ux = ux * ut; uy = uy * ut; uz = uz * ut;
  const double dutdt = D.dutdt, duxdx = D.duxdx, duydy = D.duydy, duzdz = D.duzdz;
  const double duzdt = D.duzdt, duzdx = D.duzdx, duzdy = D.duzdy, duxdt = D.duxdt;
  const double duxdy = D.duxdy, duxdz = D.duxdz, duydt = D.duydt, duydx = D.duydx, duydz = D.duydz;
  const double ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
  if (coordinates == MINKOWSKI) {
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
    const double t = M.t;
    const double t2 = pow(t, 2.), t3 = t * t * t;   // t**3. as flang lowers it (two multiplications)
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

// The EoS values of the viscous source at (rh, pr), with one pressure search
// and, off the table's nodes, no energy search: en as eos_energy returns it;
// temp, s as get_derived_data returns them, and cs2 from eos_sound at that
// function's energy (the steps of get_viscous_parameters).  Returns the
// energy error; the sound error goes to err_sound.
__device__ __forceinline__ int visc_eos(const Eos& E, double rh, double pr, double& en, double& temp,
                                        double& s, double& cs2, int& err_sound) {
  err_sound = 0;
  if (E.eqstate != 3) {
    const int err = eos_energy(E, rh, pr, en);
    if (err > 0) return err;
    double en2;
    get_derived_data(E, rh, pr, en2, temp, s);
    err_sound = eos_sound(E, rh, en2, pr, cs2);
    return err;
  }
  if (pr < E.tp[0]) return 125;
  if (pr > E.tp[E.n - 1]) return 126;
  const int l = bracket_p(E, pr);
  const double p2 = E.tp[l], p1 = E.tp[l - 1], e2 = E.te[l], e1 = E.te[l - 1];
  const double t2 = E.tt[l], t1 = E.tt[l - 1];
  en = e1 + (e2 - e1) * (pr - p1) / (p2 - p1);
  const double f = (pr - p1) / (p2 - p1);
  temp = t1 + (t2 - t1) * f;
  const double en2 = e1 + (e2 - e1) * f;
  s = (en2 + pr) / temp;
  if (en2 < E.te[0]) { err_sound = 131; return 0; }
  if (en2 > E.te[E.n - 1]) { err_sound = 132; return 0; }
  const int m = (en2 > e1 && en2 < e2) ? l : bracket_e(E, en2);
  const double c2 = E.tc[m], c1 = E.tc[m - 1], g2 = E.te[m], g1 = E.te[m - 1];
  cs2 = c1 + (c2 - c1) * (en2 - g1) / (g2 - g1);
  return 0;
}

// system_source, system.f90:1587-1842 for HD (VISC=false) and viscous HD.
// v: the cell's primitives (a private copy: pi^zz may be rewritten).
template <int SYS, bool ZZ>
__device__ __forceinline__ int sys_source(const Eos& E, const Metric& M, int coordinates,
                                          const ViscParams& P, double* v, const Deriv& D,
                                          double* src) {
  constexpr bool VISC = SysT<SYS>::VISC;
  constexpr int NV = SysT<SYS>::NV;
#pragma unroll
  for (int k = 0; k < NV; k++) src[k] = 0.;
  double rh = v[KRH], vx = v[KVX], vy = v[KVY], vz = v[KVZ], pr = v[KPR], en;
  double vtemp = 0., vs = 0., vcs2 = 0.;
  int err_sound = 0;
  int err = VISC ? visc_eos(E, rh, pr, en, vtemp, vs, vcs2, err_sound) : eos_energy(E, rh, pr, en);
  if (err > 0) return err;
  const double t = M.t;
  double v2 = M.g1 * vx * vx + M.g2 * vy * vy + M.g3 * vz * vz;
  double glf = 1. / sqrt(1. - v2);
  double ut = glf;
  if constexpr (!VISC) {
    if (coordinates == BJORKEN) {
      double ww = (en + pr) * (glf * glf);
      if constexpr (SysT<SYS>::MHD) {
        // system.f90:1643-1652
        double Ef[3];
        ideal_E(M, v, Ef);
        double B2 = v[KBX] * v[KBX] + v[KBY] * v[KBY] + v[KBZ] * v[KBZ] * M.g3;
        double E2 = Ef[0] * Ef[0] + Ef[1] * Ef[1] + Ef[2] * Ef[2] / M.g3;
        double EEM = 0.5 * (E2 + B2);
        double e3 = Ef[2] / M.g3;
        src[KPR] = -t * (ww * vz * vz - e3 * e3 - v[KBZ] * v[KBZ] + (EEM + pr) / M.g3);
      } else {
        src[KPR] = -t * (ww * vz * vz + pr / M.g3);
      }
    }
  } else {
    const double lambda0 = 1., lambda1 = 1., lambda2 = 0.;
    double sigma[6], eta, zeta, tau_pi;
    get_sigma(coordinates, M, v, D, sigma);
    {
      // get_viscous_parameters, viscous.f90:172-199, on the values of visc_eos
      const double HBAR = 0.197326;   // common.f90:47
      eta = HBAR * P.eta_over_s * vs;
      if (P.init_type == SHEAR_VISCOUS_1D) eta = P.eta_over_s;
      zeta = bulk_zeta(P, eta, vcs2, vtemp, vs);
      tau_pi = (vtemp == 0 || vs == 0) ? 0. : P.tau_pi_coeff * eta / (vtemp * vs);
      if (P.init_type == GUBSER_VISCOUS) tau_pi = 5. * eta / (vtemp * vs);
      err = err_sound;
    }
    if (err > 0) return err;
    const double tau_pi_big = tau_pi;
    PiDer q = derived_pi<ZZ>(v, M);
    const double pitx = q.pitx, pity = q.pity, pitz = q.pitz;
    const double* p = v;
    const double g1 = M.g1, g2 = M.g2, g3 = M.g3;
    double sum_der = D.dutdt + D.duxdx + D.duydy + D.duzdz;
    double sum_der_0, sum_der_x, sum_der_y, sum_der_z, sdt;
    if (coordinates == BJORKEN) {
      sdt = sum_der + ut / t;
      sum_der_0 = M.g0 * (glf * D.dutdt + glf * p[KVX] * D.dutdx + glf * p[KVY] * D.dutdy + glf * p[KVZ] * D.dutdz);
      sum_der_x = g1 * (glf * D.duxdt + glf * p[KVX] * D.duxdx + glf * p[KVY] * D.duxdy + glf * p[KVZ] * D.duxdz);
      sum_der_y = g2 * (glf * D.duydt + glf * p[KVX] * D.duydx + glf * p[KVY] * D.duydy + glf * p[KVZ] * D.duydz);
      sum_der_z = glf * (pow(t, 2.) * D.duzdt + 2. * t * p[KVZ]) + pow(t, 2.) * (glf * p[KVX] * D.duzdx + glf * p[KVY] * D.duzdy + glf * p[KVZ] * D.duzdz);
      double ww = (en + pr + p[KPIBU]) * (glf * glf);
      src[KPR] = -t * (ww * vz * vz + (pr + p[KPIBU]) / g3 + p[KPIZZ]);
      src[KPIBU] = -1. / tau_pi_big * (p[KPIBU] + zeta * sdt) - lambda0 * 4. / 3. * p[KPIBU] * sdt;
    } else {
      sdt = sum_der;
      sum_der_0 = M.g0 * (glf * D.dutdt + p[KVX] * D.dutdx + p[KVY] * D.dutdy + p[KVZ] * D.dutdz);
      sum_der_x = g1 * (glf * D.duxdt + p[KVX] * D.duxdx + p[KVY] * D.duxdy + p[KVZ] * D.duxdz);
      sum_der_y = g2 * (glf * D.duydt + p[KVX] * D.duydx + p[KVY] * D.duydy + p[KVZ] * D.duydz);
      sum_der_z = g3 * (glf * D.duzdt + p[KVX] * D.duzdx + p[KVY] * D.duzdy + p[KVZ] * D.duzdz);
      src[KPIBU] = -1. / tau_pi_big * (p[KPIBU] + zeta * sum_der) + lambda0 * (-1. / 3. * p[KPIBU] * sum_der);
    }
    double w0x = -0.5 * (D.dutdx + D.duxdt) + 0.5 * glf * (sum_der_x - p[KVX] * g1 * sum_der_0);
    double w0y = -0.5 * (D.dutdy + D.duydt) + 0.5 * glf * (sum_der_y - p[KVY] * g2 * sum_der_0);
    double w0z = -0.5 * (D.dutdz + D.duzdt) + 0.5 * glf * (sum_der_z - p[KVZ] * g3 * sum_der_0);
    double wxy = 0.5 * (g2 * D.duydx - D.duxdy) + 0.5 * glf * (g1 * p[KVX] * sum_der_y - g2 * p[KVY] * sum_der_x);
    double wxz = 0.5 * (g3 * D.duzdx - D.duxdz) + 0.5 * glf * (g1 * p[KVX] * sum_der_z - g3 * p[KVZ] * sum_der_x);
    double wyz = 0.5 * (g3 * D.duzdy - D.duydz) + 0.5 * glf * (g2 * p[KVY] * sum_der_z - g3 * p[KVZ] * sum_der_y);
    double wx0 = w0x, wy0 = w0y, wz0 = w0z, wyx = -wxy, wzx = -wxz, wzy = -wyz;
    double IVxx = 2. * (pitx * wx0 + p[KPIXY] * wxy + p[KPIXZ] * wxz);
    double IVyy = 2. * (pity * wy0 + p[KPIXY] * wyx + p[KPIYZ] * wyz);
    double IVzz = 2. * (pitz * wz0 + p[KPIXZ] * wzx + p[KPIYZ] * wzy);
    double IVxy = p[KPIXX] * wyx + p[KPIYY] * wxy + p[KPIXZ] * wyz + p[KPIYZ] * wxz + pitx * wy0 + pity * wx0;
    double IVxz = p[KPIXX] * wzx + p[KPIZZ] * wxz + p[KPIXY] * wzy + p[KPIYZ] * wxy + pitx * wz0 + pitz * wx0;
    double IVyz = p[KPIYY] * wzy + p[KPIZZ] * wyz + p[KPIXY] * wzx + p[KPIXZ] * wyx + pity * wz0 + pitz * wy0;
    if (coordinates == BJORKEN) {
      const double sum_der_t = sdt;
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
#pragma unroll
    for (int k = KPIBU; k < NV; k++) src[k] = src[k] * v[KRH];
  }
#pragma unroll
  for (int k = 0; k < NV; k++) src[k] = src[k] * M.gp;
  return 0;
}

// Sources of the resistive systems: the fluid system's (above), the
// electromagnetic stress in the Bjorken energy source (the T^{eta eta} of
// the field: EEM / g33 - (E^eta)^2 - (B^eta)^2, system.f90:1647-1652 with the
// evolved E), and in the E equations the convective current -rho_c v
// (explicit) and Ohm's conduction current (ohm_stiff, returned in sstiff,
// times sqrt(-g)); rhoc is div E at the cell.  sig, sch: the electric and
// chiral conductivities in fm^-1.
template <int SYS, bool ZZ>
__device__ __forceinline__ int res_source(const Eos& E, const Metric& M, int coordinates, const ViscParams& P,
                                          double* v, const Deriv& D, double rhoc, double sig, double sch,
                                          double* src, double* sstiff) {
  using T = SysT<SYS>;
  const int e = sys_source<T::FLUID, ZZ>(E, M, coordinates, P, v, D, src);
  if (e > 0) return e;
#pragma unroll
  for (int k = T::NF; k < T::NV; k++) src[k] = 0.;
  const double* B = &v[T::KB];
  const double* Ev = &v[T::KE];
  const double vel[3] = {v[KVX], v[KVY], v[KVZ]};
  if (coordinates == BJORKEN) {
    const EmTerms t = em_terms(M, Ev, B);
    src[KPR] = src[KPR] + (-M.t * (t.eem / M.g3 - Ev[2] * Ev[2] - B[2] * B[2])) * M.gp;
  }
  double S[3];
  ohm_stiff(M, vel, Ev, B, sig, sch, S);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    src[T::KE + k] = -rhoc * vel[k] * M.gp;
    sstiff[k] = S[k] * M.gp;
  }
  return 0;
}

}  // namespace echo
