/*
 * ORACLE - test infrastructure, not product code.
 *
 * CPU restatement of the learned con2prim (the MLP initial guess, its
 * training gradients and the FP64 Newton refinement with its residual gate)
 * that the HIP library implements.  The reference has no such component; it
 * is pinned by mathematics instead:
 *   - the backward pass by central finite differences of the FP64 forward
 *     (orc_mlp_loss with round = 0) -- tests/test_mlp_oracle.py;
 *   - the refinement by the plain solver of the reference (algorithm 5 of
 *     system_cons2prim, system.f90:945-998, restated in oracle_sys.c and
 *     bit-exact against the reference): both must reach the same root, and
 *     by prim2cons round trips (a state built from primitives is recovered);
 *   - the derivative of the refinement's residual by finite differences.
 *
 * Network: 8 -> 64 -> 64 -> 64 -> 5, activation hard-swish
 * h(x) = x * clamp(x * c6 + 1/2, 0, 1) with c6 the float nearest 1/6.  With
 * round = 1 the forward pass follows the kernel's half-precision arithmetic
 * (IEEE binary16, round to nearest even): the network input and the weights
 * are rounded to half, and each hidden activation is evaluated in half from
 * the pre-activation z as x = half(z), t = clamp(half(x * half(c6) + 1/2),
 * 0, 1) (a fused multiply-add, one rounding), h = half(x * t).  Sums are
 * accumulated in FP64 here and in FP32 on the GPU.  orc_f16 and orc_act_h
 * are pinned against numpy's float16 in tests/test_mlp_oracle.py.
 *
 * Parameter block (float, MLP_NPARAM = 9221), row-major W[out][in]:
 *   W1[64][8] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[5][64] b4[5]
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle_mlp.h"

#define NI MLP_NIN
#define NH MLP_NH
#define NO MLP_NOUT

static const int OFF_W[4] = {0, 576, 4736, 8896};
static const int OFF_B[4] = {512, 4672, 8832, 9216};
static const int DIN[4] = {NI, NH, NH, NH};
static const int DOUT[4] = {NH, NH, NH, NO};

/* loss weights of the five heads */
static const double LAMBDA[NO] = {1.0, 1.0, 0.25, 0.25, 0.25};

#define ACT_C6 0.16666667f

/* x rounded to IEEE binary16 (nearest, ties to even; subnormals down to
 * 2^-24; +-inf beyond 65504 + half an ulp), returned exactly as a double.
 * A binary16 value with exponent e (2^(e-1) <= |x| < 2^e) has an ulp of
 * 2^(e-11), never below the subnormal ulp 2^-24. */
double orc_f16(double x) {
  if (x == 0.0 || !isfinite(x)) return x;
  int e;
  frexp(x, &e);
  int q = e - 11;
  if (q < -24) q = -24;
  const double y = ldexp(nearbyint(ldexp(x, -q)), q);   /* ties to even */
  return fabs(y) > 65504.0 ? copysign(INFINITY, x) : y;
}

/* The kernel's activation in half precision (see the header).  The
 * products of two binary16 values and the sum with 1/2 are exact in
 * double, so each orc_f16 below is the single rounding of the hardware
 * operation. */
double orc_act_h(float z) {
  const double c6 = orc_f16(ACT_C6);
  const double x = orc_f16(z);
  const double t = fmin(fmax(orc_f16(x * c6 + 0.5), 0.0), 1.0);
  return orc_f16(x * t);
}

/* the activation in float with an explicit fma (the function that the
 * half-precision forward pass approximates) */
float orc_act(float x) {
  const float a = fmaf(x, ACT_C6, 0.5f);
  return x * fminf(fmaxf(a, 0.0f), 1.0f);
}

/* dh/dx: s + x c6 where the clamp is inactive, else s (0 or 1) */
float orc_act_d(float x) {
  const float a = fmaf(x, ACT_C6, 0.5f);
  const float sv = fminf(fmaxf(a, 0.0f), 1.0f);
  return (a > 0.0f && a < 1.0f) ? fmaf(x, ACT_C6, sv) : sv;
}

static double act64(double x) {
  const double a = x * (double)ACT_C6 + 0.5;
  return x * fmin(fmax(a, 0.0), 1.0);
}

static double act64_d(double x) {
  const double a = x * (double)ACT_C6 + 0.5;
  const double sv = fmin(fmax(a, 0.0), 1.0);
  return (a > 0.0 && a < 1.0) ? sv + x * (double)ACT_C6 : sv;
}

/* Invariants of a de-densitised MHD conserved state u = (D, S_i, U, B^i)
 * (the quantities system_cons2prim forms at system.f90:886-893). */
static void invariants(const double *u, double g33, double *S2, double *B2, double *SB) {
  *S2 = u[1] * u[1] + u[2] * u[2] + u[3] * u[3] / g33;
  *B2 = u[5] * u[5] + u[6] * u[6] + u[7] * u[7] * g33;
  *SB = u[1] * u[5] + u[2] * u[6] + u[3] * u[7];
}

/* Network input: 8 scale-free features of u (u already divided by sqrt(-g)).
 * With s2 = S^2/U^2, b2 = B^2/U, q = (S.B)^2/(S^2 B^2), d = D/U: the ratios
 * and 1 - s2 in FP64, the logarithms, square root and the two quotients in
 * float, as the kernel does (its hardware log2 / sqrt / reciprocal agree
 * with the float library to an ulp or two). */
void orc_mlp_features(const double *u, double g33, float *x) {
  double S2, B2, SB;
  invariants(u, g33, &S2, &B2, &SB);
  const double U = u[4];
  const double s2 = S2 / (U * U), b2 = B2 / U;
  const double den = S2 * B2;
  const double q = den > 0.0 ? (SB * SB) / den : 0.0;
  const double d = u[0] / U;
  const double oms = fmax(1.0 - s2, 1e-30);
  const float omsf = (float)oms, b2f = (float)fmax(b2, 1e-30), df = (float)fmax(d, 1e-30);
  const float s2f = (float)s2, qf = (float)q;
  x[0] = logf(omsf) / 3.0f + 1.0f;
  x[1] = 2.0f * s2f - 1.0f;
  x[2] = logf(b2f) / 4.0f + 1.0f;
  x[3] = 2.0f * b2f / (1.0f + b2f) - 1.0f;
  x[4] = 2.0f * qf - 1.0f;
  x[5] = logf(df) / 8.0f;
  x[6] = 2.0f * sqrtf(qf) - 1.0f;
  x[7] = 2.0f * omsf / (omsf + b2f) - 1.0f;
}

/* Baseline of the W head: W0 = U' + p0 with p0 the conformal hydro closed
 * form p0 = (-U' + sqrt(4 U'^2 - 3 S^2))/3 at U' = U - B^2/2 (the state with
 * its magnetic energy removed), floored at 1e-3 U'. */
double orc_mlp_w0(const double *u, double g33) {
  double S2, B2, SB;
  invariants(u, g33, &S2, &B2, &SB);
  const double Up = fmax(u[4] - 0.5 * B2, 1e-3 * u[4]);
  const double disc = fmax(4.0 * Up * Up - 3.0 * S2, 0.0);
  const double p0 = fmax((-Up + sqrt(disc)) / 3.0, 1e-3 * Up);
  return Up + p0;
}

/* Training targets (normalised) of a labelled pair (u, v), p = e/3:
 * t0 = z/3 - 1 with z = -log(1 - v^2); t1 = 25 log(W/W0) with
 * W = 4 p gamma^2 (the head the refinement uses); t2 = (log(p/U) + 4)/3;
 * t3 = 2 U/(W + B^2) - 1; t4 = 2 v^2 - 1. */
void orc_mlp_targets(const double *u, const double *v, double g33, double *t) {
  double S2, B2, SB;
  invariants(u, g33, &S2, &B2, &SB);
  const double U = u[4];
  const double v2 = v[1] * v[1] + v[2] * v[2] + v[3] * v[3] * g33;
  const double z = -log(1.0 - v2);
  const double W = 4.0 * v[4] / (1.0 - v2);
  t[0] = z / 3.0 - 1.0;
  t[1] = 25.0 * log(W / orc_mlp_w0(u, g33));
  t[2] = (log(v[4] / U) + 4.0) / 3.0;
  t[3] = 2.0 * U / (W + B2) - 1.0;
  t[4] = 2.0 * v2 - 1.0;
}

/* Forward pass of one state.  round = 1: half-precision input, weights and
 * hidden activations (orc_act_h), outputs rounded to float, as on the GPU;
 * round = 0: FP64 throughout.  zs (optional): the 3x64 hidden
 * pre-activations; hs (optional): the layer inputs h0 (8) then h1..h3 (64
 * each), as used by the next layer. */
static void forward(const float *P, const float *x, int round, double *y, double *zs, double *hs) {
  double h[NH], z[NH];
  int n_in = NI;
  for (int i = 0; i < NI; i++) h[i] = round ? orc_f16(x[i]) : x[i];
  if (hs) for (int i = 0; i < NI; i++) hs[i] = h[i];
  for (int l = 0; l < 4; l++) {
    const float *W = P + OFF_W[l], *b = P + OFF_B[l];
    for (int o = 0; o < DOUT[l]; o++) {
      double acc = b[o];
      for (int i = 0; i < n_in; i++) {
        const double w = round ? orc_f16(W[o * n_in + i]) : W[o * n_in + i];
        acc += w * h[i];
      }
      z[o] = acc;
    }
    if (l == 3) {
      for (int o = 0; o < NO; o++) y[o] = round ? (double)(float)z[o] : z[o];
      return;
    }
    for (int o = 0; o < NH; o++) {
      if (zs) zs[l * NH + o] = z[o];
      h[o] = round ? orc_act_h((float)z[o]) : act64(z[o]);
      if (hs) hs[NI + l * NH + o] = h[o];
    }
    n_in = NH;
  }
}

void orc_mlp_forward(const float *P, const float *x, int round, float *y) {
  double yy[NO];
  forward(P, x, round, yy, 0, 0);
  for (int o = 0; o < NO; o++) y[o] = (float)yy[o];
}

/* Backward pass of one state: dz holds dL/dy on entry; adds the parameter
 * gradient to grad.  round = 1 uses the half-precision weights of the
 * forward pass and the float activation derivative (the half-precision
 * roundings are passed through as the identity). */
static void backward(const float *P, int round, const double *zs, const double *hs, double *dz,
                     double *grad) {
  double dh[NH];
  for (int l = 3; l >= 0; l--) {
    const float *W = P + OFF_W[l];
    const int ni = DIN[l], no = DOUT[l];
    const double *hin = (l == 0) ? hs : hs + NI + (l - 1) * NH;
    double *gW = grad + OFF_W[l], *gb = grad + OFF_B[l];
    for (int o = 0; o < no; o++) {
      gb[o] += dz[o];
      for (int i = 0; i < ni; i++) gW[o * ni + i] += dz[o] * hin[i];
    }
    if (l == 0) return;
    for (int i = 0; i < ni; i++) {
      double acc = 0.0;
      for (int o = 0; o < no; o++)
        acc += (round ? orc_f16(W[o * ni + i]) : (double)W[o * ni + i]) * dz[o];
      dh[i] = acc;
    }
    for (int i = 0; i < ni; i++) {
      const double zz = zs[(l - 1) * NH + i];
      dz[i] = dh[i] * (round ? (double)orc_act_d((float)zz) : act64_d(zz));
    }
  }
}

/* Mean weighted squared error over n labelled states (u, v: 8 x n,
 * variable-major, u de-densitised) and, if grad != NULL, its gradient with
 * respect to the parameter block (FP64 accumulation). */
double orc_mlp_loss(const float *P, const double *u, const double *v, long n, double g33,
                    int round, double *grad) {
  double loss = 0.0;
  if (grad) memset(grad, 0, sizeof(double) * MLP_NPARAM);
  for (long s = 0; s < n; s++) {
    double us[8], vs[8], t[NO], y[NO], dz[NH], zs[3 * NH], hs[NI + 3 * NH];
    float x[NI];
    for (int k = 0; k < 8; k++) { us[k] = u[k * n + s]; vs[k] = v[k * n + s]; }
    orc_mlp_features(us, g33, x);
    orc_mlp_targets(us, vs, g33, t);
    forward(P, x, round, y, zs, hs);
    for (int o = 0; o < NO; o++) {
      const double r = y[o] - t[o];
      loss += LAMBDA[o] * r * r;
      dz[o] = 2.0 * LAMBDA[o] * r / (double)n;
    }
    if (grad) backward(P, round, zs, hs, dz, grad);
  }
  return loss / (double)n;
}

/* ---------------------------------------------------------------------- */
/* FP64 refinement for ideal MHD with p = e/3.  Algorithm 5 of the reference
 * (solver8, system.f90:2152-2201) solves two equations in (x = v^2,
 * W = (e+p) gamma^2): the energy cubic
 *   R1 = c3 W^3 + c2 W^2 + c0,  c3 = (3 + x)/4, c2 = B^2 (1 + x)/2 - U,
 *        c0 = -(S.B)^2/2,
 * nested inside rtsafe on the momentum equation
 *   R2 = (W + B^2)^2 x - (2W + B^2) (S.B)^2 / W^2 - S^2.
 * R2 is linear in x, so x(W) = (S^2 + (2W + B^2)(S.B)^2/W^2) / (W + B^2)^2
 * exactly, and the refinement is Newton on f(W) = R1(x(W), W) alone, started
 * from the network's W. */
static void e3p_w(double W, double S2, double B2, double SB2, double U, double *f, double *df,
                  double *xo, double *dxo) {
  const double iw = 1.0 / W;
  const double wb = W + B2;
  const double iwb = 1.0 / wb;
  const double sbw2 = SB2 * iw * iw;
  const double x = (S2 + (2.0 * W + B2) * sbw2) * iwb * iwb;
  const double dxdw = (2.0 * sbw2 - 2.0 * (2.0 * W + B2) * sbw2 * iw) * iwb * iwb - 2.0 * x * iwb;
  const double c3 = 0.25 * (3.0 + x);
  const double c2 = 0.5 * B2 * (1.0 + x) - U;
  *f = (c3 * W + c2) * W * W - 0.5 * SB2;
  *df = W * (3.0 * c3 * W + 2.0 * c2) + W * W * (0.25 * W + 0.5 * B2) * dxdw;
  *xo = x;
  *dxo = dxdw;
}

void orc_e3p_w(double W, double S2, double B2, double SB2, double U, double *out3) {
  double dx;
  e3p_w(W, S2, B2, SB2, U, &out3[0], &out3[1], &out3[2], &dx);
}

/* From network outputs y (normalised; y[1] = 25 log(W/W0)) to primitives.
 * Returns 0 when the state passes the gate (after `iters` Newton iterations
 * the Newton correction dW at the final W is below tol * W, the change it
 * implies in 1 - v^2 is below tol * (1 - v^2), and 0 <= v^2 < 1),
 * 1 when it must fall back to the plain solver.  u: 8 conserved values
 * multiplied by sqrt(-g) (as stored on the grid); gm = 1/sqrt(-g). */
int orc_mlp_refine(const double *u_in, const float *y, double g33, double gm, int iters,
                   double tol, double *v) {
  double u[8];
  for (int k = 0; k < 8; k++) u[k] = gm * u_in[k];
  double S2, B2, SB;
  invariants(u, g33, &S2, &B2, &SB);
  const double SB2 = SB * SB, U = u[4];
  /* the guess in float precision, as the kernel's hardware exp2 */
  double W = orc_mlp_w0(u, g33) * (double)exp2f(y[1] * (1.44269504f / 25.0f));
  double f, df, x, dxdw;
  for (int it = 0; it < iters; it++) {
    e3p_w(W, S2, B2, SB2, U, &f, &df, &x, &dxdw);
    W -= f / df;
  }
  e3p_w(W, S2, B2, SB2, U, &f, &df, &x, &dxdw);
  const double dW = f / df;
  /* the correction's effect on W and on 1 - v^2 (p and rho scale with it) */
  const int ok = (W > 0.0) && (x >= 0.0) && (x < 1.0) && (fabs(dW) <= tol * W) &&
                 (fabs(dxdw * dW) <= tol * (1.0 - x));
  if (!ok) return 1;
  /* primitives as at the end of algorithm 5 (system.f90:990-998) */
  const double vb = SB / W;
  v[0] = u[0] * sqrt(1.0 - x);
  v[4] = 0.25 * (1.0 - x) * W;
  v[1] = (u[1] + vb * u[5]) / (W + B2);
  v[2] = (u[2] + vb * u[6]) / (W + B2);
  v[3] = (u[3] / g33 + vb * u[7]) / (W + B2);
  v[5] = u[5];
  v[6] = u[6];
  v[7] = u[7];
  return 0;
}
