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
 *   - the initial magnetic field of generate_B_field_profile
 *     (glaubermc.f90:877-1174, nucleus-nucleus branch): orc_glauber_bfield.
 * The participant, collision and spectator positions are inputs (the
 * reference draws them in generate_events with its own random-number
 * stream).
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
          const double zeta = bulk_zeta(eta, cs2, temp, s);
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

/* The field of one proton of nucleus A (a = 1) or B (a = 0) at transverse
 * offset (xc, yc) from the cell and longitudinal distance zpos, scaled by
 * the prefactor w (B_amplify_factor, times pw for participants), added to
 * (bx, by, bz) as glaubermc.f90:1028-1056 (A) and 1099-1127 (B) do: the
 * classical azimuthal field Bphi and the chiral radial and longitudinal
 * fields Br, Bz of a charge moving through a medium of electric
 * conductivity sigma_el and chiral conductivity sigma_chi.  pow(D, 1.5) is
 * D**1.5 (a real exponent). */
static void proton_field(int a, double w, double xc, double yc, double zpos, double gam, double vel,
                         int magfield_type, double tom, double tomz, double *bx, double *by,
                         double *bz) {
  const double hbar = 0.197326, sqrh = 0.444214, qbig = 0.30282212;   /* common.f90:51-55 */
  const double pi = acos(-1.);
  const double sel = G.sigma_el, sch = G.sigma_chi;
  const double phi = atan2(yc, xc);
  const double xt2 = xc * xc + yc * yc;
  const double xt = sqrt(xt2);
  const double D = (gam * zpos) * (gam * zpos) + xt2;
  const double sqrD = sqrt(D);
  const double A = (sel * vel * gam / 2) * (gam * zpos - sqrD) / hbar;
  const double D15 = pow(D, 1.5);
  double bphi = 0., br = 0., bzz = 0.;
  if (magfield_type != 3)
    bphi = w * qbig / (4 * pi) * sqrh * (vel * gam * xt / D15) * (1 + sel * vel * gam * sqrD / (2 * hbar)) *
           exp(A);
  if (magfield_type != 2) {
    br = w * (-sch) * qbig / (8 * pi) * (vel * (gam * gam) * xt / D15) * (gam * zpos + A * sqrD) * exp(A) /
         sqrh;
    bzz = w * sch * qbig / (8 * pi) * (vel * gam / D15) *
          ((gam * zpos) * (gam * zpos) * (1 + sel * vel * gam * sqrD / (2 * hbar)) +
           D * (1 - sel * vel * gam * sqrD / (2 * hbar))) *
          exp(A) / sqrh;
  }
  if (a) {
    *bx = *bx + (bphi * (-sin(phi)) + br * cos(phi)) / tom;
    *by = *by + (bphi * cos(phi) + br * sin(phi)) / tom;
    *bz = *bz + bzz / tomz;
  } else {
    *bx = *bx + (bphi * sin(phi) + br * cos(phi)) / tom;
    *by = *by + (bphi * (-cos(phi)) + br * sin(phi)) / tom;
    *bz = *bz + (-bzz) / tomz;
  }
}

/* generate_B_field_profile, glaubermc.f90:1009-1168: adds the field of the
 * protons of the spectators and participants of nuclei A and B to B^x, B^y,
 * B^eta of the primitives v (nv x cells, variable-major), cell by cell and
 * charge by charge in the reference's order (spectators A, participants A,
 * spectators B, participants B).  Each list is read up to its first
 * neutron (type != 0), as the reference does.  par: gamma_col,
 * B_amplify_factor, pw, magfield_type (1 classical + chiral, 2 classical,
 * 3 chiral).  Returns 1 when the grid has no magnetic field. */
int orc_glauber_bfield(const double *par, const double *sa, const int *ta, int nsa, const double *pa,
                       const int *tpa, int npa, const double *sb, const int *tb, int nsb,
                       const double *pb, const int *tpb, int npb, double *v) {
  if (!G.mhd) return 1;
  const double gam = par[0], amp = par[1], pw = par[2];
  const int magfield_type = (int)par[3];
  const double vel = sqrt(1. - 1. / (gam * gam));
  const size_t n = (size_t)G.nx * G.ny * G.nz;
  const double *xy[4] = {sa, pa, sb, pb};
  const int *ty[4] = {ta, tpa, tb, tpb};
  const int cnt[4] = {nsa, npa, nsb, npb};
  const double wt[4] = {amp, amp * pw, amp, amp * pw};
  for (int iz = 0; iz < G.nz; iz++) {
    double tt, zcoord, tom, tomz;
    if (G.coordinates == 1) {
      zcoord = G.zc[iz]; tt = G.tmin; tom = 1.; tomz = 1.;
    } else {
      tt = G.tmin * cosh(G.zc[iz]); zcoord = G.tmin * sinh(G.zc[iz]);
      tom = cosh(G.zc[iz]); tomz = G.tmin;
    }
    for (int iy = 0; iy < G.ny; iy++)
      for (int ix = 0; ix < G.nx; ix++) {
        const size_t c = ((size_t)iz * G.ny + iy) * G.nx + ix;
        double bx = v[G.kbx * n + c], by = v[G.kby * n + c], bz = v[G.kbz * n + c];
        for (int l = 0; l < 4; l++) {
          const double zpos = l < 2 ? tt * vel + zcoord : tt * vel - zcoord;
          for (int i = 0; i < cnt[l] && ty[l][i] == 0; i++)
            proton_field(l < 2, wt[l], G.xc[ix] - xy[l][2 * i], G.yc[iy] - xy[l][2 * i + 1], zpos, gam,
                         vel, magfield_type, tom, tomz, &bx, &by, &bz);
        }
        v[G.kbx * n + c] = bx;
        v[G.kby * n + c] = by;
        v[G.kbz * n + c] = bz;
      }
  }
  return 0;
}
