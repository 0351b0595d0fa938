/*
 * ORACLE - test infrastructure, not product code.
 *
 * Glauber Monte-Carlo initial state, restated for checking the GPU
 * generator:
 *   - the energy deposition of generate_energy_density_profile
 *     (glaubermc.f90:686-842, nucleus-nucleus branch): Gaussian sources at
 *     the participants of nuclei A and B (weighted (ybeam +- eta)/ybeam) and
 *     at the binary collisions (fraction ah), shifted to the centre of mass
 *     of the sources, times the longitudinal profile feta
 *     (glaubermc.f90:661-682);
 *   - the primitive state of the Glauber-MC initialisation (init.f90:
 *     2662-2730): rho = 1, v = 0, p = eos_pressure(rhocenter, e + enezero);
 *   - viscous runs: the Navier-Stokes initial stresses of viscous_initio
 *     (viscous.f90:101-136) with get_viscous_parameters_0 (140-167).
 * The participant and collision positions are inputs (the reference draws
 * them in generate_events with its own random-number stream).
 */
#include <math.h>
#include <string.h>

#include "oracle.h"
#include "oracle_int.h"

static const double HBAR = 0.197326;   /* common.f90:51 */

static double feta(double etas, double ybeam, double deta, double sigeta) {
  const double half_deta = deta / 2.;
  if (fabs(etas) >= ybeam) return 0.;
  if (fabs(etas) <= half_deta) return 1.;
  const double a = fabs(etas) - half_deta;
  return exp(-(a * a) / (2. * (sigeta * sigeta)));
}

/* par: kappa, ah, sig_mc, ybeam, deta, sigeta, rhocenter.
 * pa, pb, pc: (x, y) pairs of the participants of A, B and the collisions.
 * vout: the primitive state (nv x cells, Fortran layout). */
int orc_glauber_prim(const double *par, const double *pa, int na, const double *pb, int nb,
                     const double *pc, int nc, double *vout) {
  const double kappa = par[0], ah = par[1], sig = par[2], ybeam = par[3], deta = par[4];
  const double sigeta = par[5], rhocenter = par[6];
  const double pi = acos(-1.);
  double cmx_p = 0., cmy_p = 0., cmx_c = 0., cmy_c = 0.;
  for (int i = 0; i < na; i++) { cmx_p = cmx_p + pa[2 * i]; cmy_p = cmy_p + pa[2 * i + 1]; }
  for (int i = 0; i < nb; i++) { cmx_p = cmx_p + pb[2 * i]; cmy_p = cmy_p + pb[2 * i + 1]; }
  for (int i = 0; i < nc; i++) { cmx_c = cmx_c + pc[2 * i]; cmy_c = cmy_c + pc[2 * i + 1]; }
  const double den = (1 - ah) * (na + nb) + ah * nc;
  const double cmx = ((1 - ah) * cmx_p + ah * cmx_c) / den;
  const double cmy = ((1 - ah) * cmy_p + ah * cmy_c) / den;
  const double s2 = sig * sig;
  const size_t n = (size_t)G.nx * G.ny * G.nz;
  memset(vout, 0, sizeof(double) * n * G.nv);
  for (int iz = 0; iz < G.nz; iz++)
    for (int iy = 0; iy < G.ny; iy++)
      for (int ix = 0; ix < G.nx; ix++) {
        const double x = G.xc[ix], y = G.yc[iy], z = G.zc[iz];
        double ea = 0., eb = 0., ec = 0.;
        for (int i = 0; i < na; i++) {
          const double a = x - pa[2 * i] + cmx, b = y - pa[2 * i + 1] + cmy;
          ea = ea + exp(-(a * a + b * b) / 2. / s2);
        }
        for (int i = 0; i < nb; i++) {
          const double a = x - pb[2 * i] + cmx, b = y - pb[2 * i + 1] + cmy;
          eb = eb + exp(-(a * a + b * b) / 2. / s2);
        }
        const double ep = ea * (ybeam + z) / ybeam + eb * (ybeam - z) / ybeam;
        for (int i = 0; i < nc; i++) {
          const double a = x - pc[2 * i] + cmx, b = y - pc[2 * i + 1] + cmy;
          ec = ec + exp(-(a * a + b * b) / 2. / s2);
        }
        const double e = kappa / (2. * pi * s2) * (ah * ec + (1 - ah) * ep) * feta(z, ybeam, deta, sigeta);
        double pr, d1, d2;
        int err = eos_pressure(rhocenter, e + G.enezero, &pr, &d1, &d2);
        if (err > 0) return err;
        const size_t c = ((size_t)iz * G.ny + iy) * G.nx + ix;
        vout[KRH * n + c] = 1.;
        vout[KPR * n + c] = pr;
        if (G.viscosity) {
          /* viscous_initio with get_viscous_parameters_0 */
          double en, temp, s, cs2;
          err = get_derived_data(1., pr, &en, &temp, &s);
          if (err > 0) return err;
          err = eos_sound(1., en, pr, &cs2);
          if (err > 0) return err;
          const double eta = HBAR * G.eta_over_s * s;
          const double zeta = G.bulkvis ? 2. * eta * (1. / 3. - cs2) : 0.;
          const double tau0 = G.tmin;
          const double g3 = G.coordinates == 2 ? tau0 * tau0 : 1.;
          vout[KPIZZ * n + c] = -g3 * 4. * eta / (3. * tau0);
          vout[KPIXX * n + c] = 2. * eta / (3. * tau0);
          vout[KPIYY * n + c] = 2. * eta / (3. * tau0);
          vout[KPIBU * n + c] = -zeta / tau0;
        }
      }
  return 0;
}
