// Learned con2prim for ideal MHD with p = e/3: an 8 -> 64 -> 64 -> 64 -> 5
// network evaluated on MFMA (half-precision operands, FP32 accumulation) gives the
// initial guess of the inversion (W = (e+p) gamma^2, with v^2 following
// from W exactly); FP64 Newton iterations refine it and a residual gate
// sends every state it does not accept to the plain solver (algorithm 5 of system_cons2prim,
// system.f90:945-998).  The network is trained on the device by the
// kernels below (forward with saved activations, backward, weight gradient,
// Adam).
//
// MFMA formulation: weights are the A operand and activations the B operand,
// D = W * H with the output features on the rows and 16 states on the
// columns.  A 16x16 D block leaves lane (c = lane % 16, g = lane / 16) with
// rows 4g..4g+3 of state c; the B operand of the 16x16x32 instruction holds
// k = 8g..8g+7 of state c.  Concatenating the D blocks 2s and 2s+1 of the
// previous layer therefore gives the B operand of K-step s directly, for the
// K order (32s + 4g + m, m < 4; 32s + 16 + 4g + m - 4, m >= 4) in which the
// A operand (the weights) is packed.  No activation leaves the registers
// between layers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "phys.hpp"

namespace echo {
namespace mlp {

constexpr int NI = 8, NH = 64, NO = 5, NPARAM = 9221;
constexpr int OFF_W1 = 0, OFF_B1 = 512, OFF_W2 = 576, OFF_B2 = 4672, OFF_W3 = 4736, OFF_B3 = 8832,
              OFF_W4 = 8896, OFF_B4 = 9216;
// packed half-precision A operands: layer 1 4 x 64 x 4, layers 2-3 8 x 64 x 8, layer 4 2 x 64 x 8
constexpr int PK_L1 = 0, PK_L2 = 1024, PK_L3 = 5120, PK_L4 = 9216, PK_N = 10240;
// biases as the accumulator initial values: 64, 64, 64, 16 (rows 5..15 zero)
constexpr int BS_L1 = 0, BS_L2 = 64, BS_L3 = 128, BS_L4 = 192, BS_N = 208;
constexpr int BLOCK = 256;            // 4 waves, each 64 states per tile
constexpr int BWD_BLOCK = 128;

typedef _Float16 h16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 h16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// activation: hard-swish h(x) = x * clamp(x / 6 + 1/2, 0, 1) -- 0 below -3,
// x above 3, x^2/6 + x/2 between.  act and act_d are the FP32 function and
// its derivative (training backward pass).  The forward pass evaluates it in
// half precision on pairs of values, as the operands of the next MFMA:
// x = half(z); t = clamp(fma(x, half(c6), 1/2), 0, 1) (one rounding, the
// clamp bit of the packed FMA); h = half(x * t).  That is one packed
// conversion, one packed FMA and one packed multiply per two values (1.5
// VALU instructions per activation; 2.5 in FP32 with a bfloat16 conversion).
// The network is VALU-bound on its activations (192 per state against
// 18 kflop of MFMA work).
#define ACT_C6 0.16666667f

__device__ __forceinline__ float act(float x) {
  return x * __builtin_amdgcn_fmed3f(__builtin_fmaf(x, ACT_C6, 0.5f), 0.0f, 1.0f);
}

// dh/dx: s + x/6 inside (-3, 3), else s (0 below, 1 above)
__device__ __forceinline__ float act_d(float x) {
  const float a = __builtin_fmaf(x, ACT_C6, 0.5f);
  const float sv = __builtin_amdgcn_fmed3f(a, 0.0f, 1.0f);
  return (a > 0.0f && a < 1.0f) ? __builtin_fmaf(x, ACT_C6, sv) : sv;
}

__device__ __forceinline__ float h16round(float x) { return (float)(_Float16)x; }

// The per-tile LDS exchanges of k_mlp (sX[w], sY[w]) stay inside one wave.
// LDS operations of a wave complete in the order it issues them, so only the
// compiler must keep each store before the loads that follow: wave-scope
// fences, no barrier across the block.  The block's waves then run out of
// step and hide one another's latencies (the FP64 refinement of one wave
// beside the loads and MFMA chains of another).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// two activations in half precision (see above)
__device__ __forceinline__ h16x2 act2h(f32x2 z) {
  const h16x2 x = __builtin_convertvector(z, h16x2);
  const h16x2 c6 = {(_Float16)ACT_C6, (_Float16)ACT_C6};
  const h16x2 hf = {(_Float16)0.5f, (_Float16)0.5f};
  h16x2 t = __builtin_elementwise_fma(x, c6, hf);
  t = __builtin_elementwise_min(__builtin_elementwise_max(t, (h16x2){0, 0}), (h16x2){1, 1});
  return x * t;
}

__device__ __forceinline__ h16x4 act4(f32x4 a) {
  const h16x2 lo = act2h((f32x2){a[0], a[1]});
  const h16x2 hi = act2h((f32x2){a[2], a[3]});
  return (h16x4){lo[0], lo[1], hi[0], hi[1]};
}

__device__ __forceinline__ h16x8 cat(h16x4 a, h16x4 b) {
  return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7);
}

__device__ __forceinline__ f32x4 mfma32(h16x8 a, h16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x4 mfma16(h16x4 a, h16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// Network inputs of one de-densitised state (u[0..7] = D, S_i, U, B^i),
// 8 scale-free features of S^2/U^2, B^2/U, (S.B)^2/(S^2 B^2) and D/U.  The
// ratios and 1 - S^2/U^2 (which cancels as v -> 1) are formed in FP64; the
// logarithms, the square root and the two quotients that follow only feed
// half-precision inputs and are evaluated in FP32 with the hardware log2,
// sqrt and reciprocal (the FP64 library calls took a third of the kernel).
__device__ __forceinline__ void features(const double* u, double g33, float* x) {
  const double S2 = u[1] * u[1] + u[2] * u[2] + u[3] * u[3] / g33;
  const double B2 = u[5] * u[5] + u[6] * u[6] + u[7] * u[7] * g33;
  const double SB = u[1] * u[5] + u[2] * u[6] + u[3] * u[7];
  const double iU = rcp_nr(u[4]);
  const double s2 = S2 * (iU * iU), b2 = B2 * iU;
  const double den = S2 * B2;
  const double q = den > 0.0 ? (SB * SB) * rcp_nr(den) : 0.0;
  const double d = u[0] * iU;
  const double oms = dmax(1.0 - s2, 1e-30);
  const float LN2 = 0.69314718f;
  const float omsf = (float)oms, b2f = (float)dmax(b2, 1e-30), df = (float)dmax(d, 1e-30);
  const float s2f = (float)s2, qf = (float)q;
  x[0] = __builtin_amdgcn_logf(omsf) * (LN2 / 3.0f) + 1.0f;
  x[1] = 2.0f * s2f - 1.0f;
  x[2] = __builtin_amdgcn_logf(b2f) * (LN2 / 4.0f) + 1.0f;
  x[3] = 2.0f * b2f * __builtin_amdgcn_rcpf(1.0f + b2f) - 1.0f;
  x[4] = 2.0f * qf - 1.0f;
  x[5] = __builtin_amdgcn_logf(df) * (LN2 / 8.0f);
  x[6] = 2.0f * __builtin_amdgcn_sqrtf(qf) - 1.0f;
  x[7] = 2.0f * omsf * __builtin_amdgcn_rcpf(omsf + b2f) - 1.0f;
}

// Baseline of the W head: W0 = U' + p0, p0 the conformal hydro closed form
// (-U' + sqrt(4 U'^2 - 3 S^2))/3 at U' = U - B^2/2 (magnetic energy
// removed), floored at 1e-3 U'.  The head predicts 25 log(W/W0), a quantity
// ~45x narrower than log(W/U) over the training states.
__device__ __forceinline__ double w_baseline(double S2, double B2, double U) {
  const double Up = dmax(U - 0.5 * B2, 1e-3 * U);
  const double disc = dmax(4.0 * Up * Up - 3.0 * S2, 0.0);
  const double p0 = dmax((-Up + sqrt(disc)) / 3.0, 1e-3 * Up);
  return Up + p0;
}

// Normalised training targets of a labelled pair (de-densitised u, v).
__device__ __forceinline__ void targets(const double* u, const double* v, double g33, double* t) {
  const double S2 = u[1] * u[1] + u[2] * u[2] + u[3] * u[3] / g33;
  const double B2 = u[5] * u[5] + u[6] * u[6] + u[7] * u[7] * g33;
  const double U = u[4];
  const double v2 = v[1] * v[1] + v[2] * v[2] + v[3] * v[3] * g33;
  const double z = -log(1.0 - v2);
  const double W = 4.0 * v[4] / (1.0 - v2);
  t[0] = z / 3.0 - 1.0;
  t[1] = 25.0 * log(W / w_baseline(S2, B2, U));
  t[2] = (log(v[4] / U) + 4.0) / 3.0;
  t[3] = 2.0 * U / (W + B2) - 1.0;
  t[4] = 2.0 * v2 - 1.0;
}

// Refinement for ideal MHD with p = e/3.  Algorithm 5 of the reference
// (solver8, system.f90:2152-2201) solves the energy cubic in W nested inside
// rtsafe on the momentum equation in x = v^2.  The momentum equation is
// linear in x, so x(W) = (S^2 + (2W + B^2)(S.B)^2/W^2) / (W + B^2)^2 exactly,
// and the refinement is Newton on f(W) = c3(x) W^3 + c2(x) W^2 - (S.B)^2/2
// alone (c3 = (3 + x)/4, c2 = B^2 (1 + x)/2 - U), from the network's W.
__device__ __forceinline__ void e3p_w(double W, double S2, double B2, double SB2, double U, double& f,
                                      double& df, double& x, double& dxdw) {
  const double iw = rcp_nr(W);
  const double wb = W + B2;
  const double iwb = rcp_nr(wb);
  const double sbw2 = SB2 * iw * iw;
  x = (S2 + (2.0 * W + B2) * sbw2) * iwb * iwb;
  dxdw = (2.0 * sbw2 - 2.0 * (2.0 * W + B2) * sbw2 * iw) * iwb * iwb - 2.0 * x * iwb;
  const double c3 = 0.25 * (3.0 + x);
  const double c2 = 0.5 * B2 * (1.0 + x) - U;
  f = (c3 * W + c2) * W * W - 0.5 * SB2;
  df = W * (3.0 * c3 * W + 2.0 * c2) + W * W * (0.25 * W + 0.5 * B2) * dxdw;
}

// Newton refinement of the network's W (y1 = 25 log(W/W0)) and the residual
// gate: accepted when the Newton correction dW at the final W is below
// tol * W, the change it implies in 1 - v^2 is below tol * (1 - v^2), and
// 0 <= v^2 < 1.  u: the de-densitised state; writes the 8 primitives.
__device__ __forceinline__ bool refine(const double* u, float y1, double g33, int iters, double tol,
                                       double* v) {
  const double S2 = u[1] * u[1] + u[2] * u[2] + u[3] * u[3] / g33;
  const double B2 = u[5] * u[5] + u[6] * u[6] + u[7] * u[7] * g33;
  const double SB = u[1] * u[5] + u[2] * u[6] + u[3] * u[7];
  const double SB2 = SB * SB, U = u[4];
  // the guess needs no more than float precision: exp2 on the hardware
  double W = w_baseline(S2, B2, U) * (double)__builtin_amdgcn_exp2f(y1 * (1.44269504f / 25.0f));
  double f, df, x, dxdw;
  for (int it = 0; it < iters; it++) {
    e3p_w(W, S2, B2, SB2, U, f, df, x, dxdw);
    W -= f * rcp_nr(df);
  }
  e3p_w(W, S2, B2, SB2, U, f, df, x, dxdw);
  const double dW = f * rcp_nr(df);
  // the correction's effect on W and on 1 - v^2 (p and rho scale with it)
  const bool ok = (W > 0.0) && (x >= 0.0) && (x < 1.0) && (fabs(dW) <= tol * W) &&
                  (fabs(dxdw * dW) <= tol * (1.0 - x));
  if (!ok) return false;
  // primitives as at the end of algorithm 5 (system.f90:990-998)
  const double vb = SB * rcp_nr(W);
  const double iwb = rcp_nr(W + B2);
  v[0] = u[0] * sqrt(1.0 - x);
  v[4] = 0.25 * (1.0 - x) * W;
  v[1] = (u[1] + vb * u[5]) * iwb;
  v[2] = (u[2] + vb * u[6]) * iwb;
  v[3] = (u[3] / g33 + vb * u[7]) * iwb;
  v[5] = u[5];
  v[6] = u[6];
  v[7] = u[7];
  return true;
}

enum { MODE_C2P = 0, MODE_FWD = 1, MODE_TRAIN = 2 };

struct NetArgs {
  const _Float16* packed;   // PK_N half-precision A operands
  const float* bias;      // BS_N
};

struct C2pArgs {
  int coordinates;
  int iters;
  double tol;
  const DevScalars* S;    // time (metric)
  int* fail_count;
  int* fail_list;
  int in_step;            // inside a grid's step: nothing to do once the run stopped
};

// Launch bounds (BLOCK, 3): three waves per SIMD (at most 168 VGPRs; the
// inference variant takes 160 with no spills).  With two, the FP64 Newton
// steps and the MFMA chains stall with too little to switch to.
// One kernel for inference (MODE_C2P), plain forward (MODE_FWD, outputs y)
// and the training forward (MODE_TRAIN: y, network inputs h0, pre-activations
// z1..z3 and rounded activations h1..h3, all [feature][state]).
// Ub: conserved states [8 or 9][stride] starting at state `off`; n states.
template <int MODE, int NVU>
__global__ __launch_bounds__(BLOCK, 3) void k_mlp(NetArgs net, C2pArgs ca, const double* __restrict__ Ub,
                                               long long stride, long long off, long long n,
                                               double* __restrict__ Vb, int* __restrict__ status,
                                               float* __restrict__ ybuf, float* __restrict__ xbuf,
                                               float* __restrict__ zbuf, float* __restrict__ hbuf) {
  __shared__ h16x4 sA1[4 * 64];
  __shared__ h16x8 sA2[8 * 64];
  __shared__ h16x8 sA3[8 * 64];
  __shared__ h16x8 sA4[2 * 64];
  __shared__ float sBias[BS_N];
  __shared__ h16x8 sX[4][64];
  __shared__ float2 sY[4][64];

  if (MODE == MODE_C2P && ca.in_step && (ca.S->run_crashed || ca.S->temp_check)) return;   // block-uniform
  {  // weights to LDS
    const uint4* src = reinterpret_cast<const uint4*>(net.packed);
    uint4* d1 = reinterpret_cast<uint4*>(sA1);
    uint4* d2 = reinterpret_cast<uint4*>(sA2);
    uint4* d3 = reinterpret_cast<uint4*>(sA3);
    uint4* d4 = reinterpret_cast<uint4*>(sA4);
    for (int i = threadIdx.x; i < PK_N / 8; i += BLOCK) {
      const uint4 w = src[i];
      if (i < PK_L2 / 8) d1[i] = w;
      else if (i < PK_L3 / 8) d2[i - PK_L2 / 8] = w;
      else if (i < PK_L4 / 8) d3[i - PK_L3 / 8] = w;
      else d4[i - PK_L4 / 8] = w;
    }
    for (int i = threadIdx.x; i < BS_N; i += BLOCK) sBias[i] = net.bias[i];
  }
  __syncthreads();

  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c = lane & 15, g = lane >> 4;
  const long long ntiles = (n + 63) / 64;
  double g33 = 1.0, gm = 1.0;
  if (MODE == MODE_C2P) {
    const Metric M = make_metric(ca.coordinates, ca.S->t);
    g33 = M.g3;
    gm = M.gm;
  }

  for (long long t0 = (long long)blockIdx.x * 4; t0 < ntiles; t0 += (long long)gridDim.x * 4) {
    const long long tile = t0 + w;
    const long long s = tile * 64 + lane;
    const bool valid = (tile < ntiles) && (s < n);
    double u[8];
    if (valid) {
#pragma unroll
      for (int k = 0; k < 8; k++) u[k] = gm * Ub[k * stride + off + s];
    } else {
      u[0] = 1.0; u[1] = 0.0; u[2] = 0.0; u[3] = 0.0; u[4] = 1.0; u[5] = 0.0; u[6] = 0.0; u[7] = 0.0;
    }
    {
      float x[8];
      features(u, g33, x);
      h16x8 xb;
#pragma unroll
      for (int k = 0; k < 8; k++) xb[k] = (_Float16)x[k];
      sX[w][lane] = xb;
      if (MODE == MODE_TRAIN && valid) {
#pragma unroll
        for (int k = 0; k < 8; k++) xbuf[k * n + s] = (float)xb[k];
      }
    }
    wave_sync();

    // layer 1 (K = 8, padded to 16)
    h16x4 H[4][4];
    f32x4 acc[4][4];
    {
      h16x4 B0[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const h16x8 xv = sX[w][16 * q + c];
        h16x4 lo = __builtin_shufflevector(xv, xv, 0, 1, 2, 3);
        h16x4 hi = __builtin_shufflevector(xv, xv, 4, 5, 6, 7);
        const h16x4 zero = {(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};
        B0[q] = g == 0 ? lo : (g == 1 ? hi : zero);
      }
#pragma unroll
      for (int rb = 0; rb < 4; rb++) {
        const f32x4 b = *reinterpret_cast<const f32x4*>(&sBias[BS_L1 + 16 * rb + 4 * g]);
        const h16x4 a = sA1[rb * 64 + lane];
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q][rb] = mfma16(a, B0[q], b);
      }
    }
#define SAVE_LAYER(L)                                                                       \
  if (MODE == MODE_TRAIN) {                                                                 \
    _Pragma("unroll") for (int q = 0; q < 4; q++) {                                         \
      const long long sq = tile * 64 + 16 * q + c;                                          \
      if (tile < ntiles && sq < n) {                                                        \
        _Pragma("unroll") for (int rb = 0; rb < 4; rb++) {                                  \
          _Pragma("unroll") for (int j = 0; j < 4; j++) {                                   \
            const int f = 16 * rb + 4 * g + j;                                              \
            zbuf[((long long)(L) * NH + f) * n + sq] = acc[q][rb][j];                       \
            hbuf[((long long)(L) * NH + f) * n + sq] = (float)H[q][rb][j];                  \
          }                                                                                 \
        }                                                                                   \
      }                                                                                     \
    }                                                                                       \
  }
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int rb = 0; rb < 4; rb++) H[q][rb] = act4(acc[q][rb]);
    SAVE_LAYER(0)

    // layers 2 and 3 (64 -> 64)
#pragma unroll
    for (int L = 1; L <= 2; L++) {
      const h16x8* sA = (L == 1) ? sA2 : sA3;
      const int bo = (L == 1) ? BS_L2 : BS_L3;
#pragma unroll
      for (int rb = 0; rb < 4; rb++) {
        const f32x4 b = *reinterpret_cast<const f32x4*>(&sBias[bo + 16 * rb + 4 * g]);
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q][rb] = b;
      }
#pragma unroll
      for (int ks = 0; ks < 2; ks++) {
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
          const h16x8 a = sA[(rb * 2 + ks) * 64 + lane];
#pragma unroll
          for (int q = 0; q < 4; q++) acc[q][rb] = mfma32(a, cat(H[q][2 * ks], H[q][2 * ks + 1]), acc[q][rb]);
        }
      }
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int rb = 0; rb < 4; rb++) H[q][rb] = act4(acc[q][rb]);
      SAVE_LAYER(L)
    }
#undef SAVE_LAYER

    // layer 4 (64 -> 5, one 16-row block)
    f32x4 y4[4];
    {
      const f32x4 b = *reinterpret_cast<const f32x4*>(&sBias[BS_L4 + 4 * g]);
#pragma unroll
      for (int q = 0; q < 4; q++) y4[q] = b;
#pragma unroll
      for (int ks = 0; ks < 2; ks++) {
        const h16x8 a = sA4[ks * 64 + lane];
#pragma unroll
        for (int q = 0; q < 4; q++) y4[q] = mfma32(a, cat(H[q][2 * ks], H[q][2 * ks + 1]), y4[q]);
      }
    }

    if (MODE != MODE_C2P) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const long long sq = tile * 64 + 16 * q + c;
        if (tile < ntiles && sq < n) {
          if (g == 0) {
#pragma unroll
            for (int j = 0; j < 4; j++) ybuf[j * n + sq] = y4[q][j];
          } else if (g == 1) {
            ybuf[4 * n + sq] = y4[q][0];
          }
        }
      }
      wave_sync();
      continue;
    }

#pragma unroll
    for (int q = 0; q < 4; q++)
      if (g == 0) sY[w][16 * q + c] = make_float2(y4[q][0], y4[q][1]);
    wave_sync();
    const float y1 = sY[w][lane].y;
    double v[8];
    const bool ok = valid && refine(u, y1, g33, ca.iters, ca.tol, v);
    if (ok) {
#pragma unroll
      for (int k = 0; k < 8; k++) Vb[k * stride + off + s] = v[k];
      if (NVU == 9) Vb[8 * stride + off + s] = Ub[8 * stride + off + s];
      if (status) status[off + s] = 0;
    }
    const bool bad = valid && !ok;
    const unsigned long long mask = __ballot(bad);
    if (mask) {
      const int cnt = __popcll(mask);
      const int lead = __ffsll((long long)mask) - 1;
      int base = 0;
      if (lane == lead) base = atomicAdd(ca.fail_count, cnt);
      base = __shfl(base, lead);
      if (bad) {
        const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0));
        ca.fail_list[base + rank] = (int)(off + s);
      }
    }
    wave_sync();
  }
}

// Plain solver on the states the gate rejected (list from k_mlp).
template <int SYS>
__global__ __launch_bounds__(256) void k_c2p_fallback(GridArgs a, const double* __restrict__ Ub,
                                                      double* __restrict__ Vb, int* __restrict__ status,
                                                      long long stride, const int* __restrict__ list,
                                                      const int* __restrict__ count) {
  constexpr int NV = SysT<SYS>::NV;
  const int m = *count;
  const Metric M = make_metric(a.coordinates, a.S->t);
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < m; i += (long long)gridDim.x * 256) {
    const long long s = list[i];
    double u[NV], v[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) { u[k] = Ub[k * stride + s]; v[k] = Vb[k * stride + s]; }
    status[s] = cons2prim<SYS, false>(a.eos, M, u, v, a.algo_solver);
#pragma unroll
    for (int k = 0; k < NV; k++) Vb[k * stride + s] = v[k];
  }
}

// The plain solver inside a grid's step, on the states the gate rejected:
// con2prim from the primitives V holds (the stage start, as in k_update),
// failures marked for k_fixup as k_update marks them.  Block 0 adds the
// number of listed states to the grid's fallback count.
template <int SYS>
__global__ __launch_bounds__(256) void k_c2p_fallback_step(GridArgs a, const double* __restrict__ U,
                                                           double* __restrict__ V, FixBufs fb,
                                                           const int* __restrict__ list,
                                                           const int* __restrict__ count) {
  constexpr int NV = SysT<SYS>::NV;
  DevScalars* S = a.S;
  if (S->run_crashed || S->temp_check) return;
  const int m = *count;
  if (blockIdx.x == 0 && threadIdx.x == 0) S->nfallback += m;
  const Metric M = make_metric(a.coordinates, S->t);
  const long long N = a.N;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < m; i += (long long)gridDim.x * 256) {
    const long long c = list[i];
    double u[NV], v[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) { u[k] = U[k * N + c]; v[k] = V[k * N + c]; }
    const int e = cons2prim<SYS, false>(a.eos, M, u, v, a.algo_solver);
    if (e != 0) {
      atomicOr(&fb.failbits[c >> 5], 1u << (c & 31));
      atomicAdd(&S->nfail, 1);
    }
#pragma unroll
    for (int k = 0; k < NV; k++) V[k * N + c] = v[k];
  }
}

static __global__ void k_zero_int(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Backward pass, one thread per state: dL/dz of the four layers
// ([feature][state]) and the loss sum.  Wq: the half-precision weights
// (as float, row-major [out][in]) that the forward pass used.
static __global__ __launch_bounds__(BWD_BLOCK) void k_mlp_bwd(const float* __restrict__ Wq, const double* __restrict__ Ub,
                                                       const double* __restrict__ Vl, long long stride,
                                                       long long off, long long n,
                                                       const float* __restrict__ ybuf,
                                                       const float* __restrict__ zbuf,
                                                       float* __restrict__ dzbuf, double* __restrict__ loss) {
  const long long s = (long long)blockIdx.x * BWD_BLOCK + threadIdx.x;
  double part = 0.0;
  if (s < n) {
    double u[8], v[8], t[NO];
#pragma unroll
    for (int k = 0; k < 8; k++) { u[k] = Ub[k * stride + off + s]; v[k] = Vl[k * stride + off + s]; }
    targets(u, v, 1.0, t);
    const float lam[NO] = {1.0f, 1.0f, 0.25f, 0.25f, 0.25f};
    float dz[NH], dh[NH];
#pragma unroll
    for (int o = 0; o < NO; o++) {
      const double r = (double)ybuf[o * n + s] - t[o];
      part += lam[o] * r * r;
      dz[o] = (float)(2.0 * lam[o] * r / (double)n);
      dzbuf[(3LL * NH + o) * n + s] = dz[o];
    }
    // layer 4 -> 3
#pragma unroll
    for (int i = 0; i < NH; i++) {
      float a = 0.0f;
#pragma unroll
      for (int o = 0; o < NO; o++) a = __builtin_fmaf(Wq[OFF_W4 + o * NH + i], dz[o], a);
      dh[i] = a;
    }
#pragma unroll
    for (int L = 2; L >= 0; L--) {
#pragma unroll
      for (int i = 0; i < NH; i++) {
        dz[i] = dh[i] * act_d(zbuf[((long long)L * NH + i) * n + s]);
        dzbuf[((long long)L * NH + i) * n + s] = dz[i];
      }
      if (L == 0) break;
      const float* W = Wq + (L == 2 ? OFF_W3 : OFF_W2);
#pragma unroll
      for (int i = 0; i < NH; i++) dh[i] = 0.0f;
#pragma unroll
      for (int o = 0; o < NH; o++) {
        const float d = dz[o];
#pragma unroll
        for (int i = 0; i < NH; i++) dh[i] = __builtin_fmaf(W[o * NH + i], d, dh[i]);
      }
    }
  }
  // block sum of the loss
  __shared__ double red[BWD_BLOCK / 64];
  for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int i = 0; i < BWD_BLOCK / 64; i++) tot += red[i];
    atomicAdd(loss, tot / (double)n);
  }
}

// Weight gradient: block b handles one output row o of one layer and sums
// dz[o][s] * h_in[i][s] over the batch in a fixed order.
static __global__ __launch_bounds__(256) void k_mlp_wgrad(const float* __restrict__ xbuf, const float* __restrict__ hbuf,
                                                   const float* __restrict__ dzbuf, long long n,
                                                   float* __restrict__ grad) {
  int b = blockIdx.x, L, o;
  if (b < 64) { L = 0; o = b; }
  else if (b < 128) { L = 1; o = b - 64; }
  else if (b < 192) { L = 2; o = b - 128; }
  else { L = 3; o = b - 192; }
  const int nin = (L == 0) ? NI : NH;
  const float* hin = (L == 0) ? xbuf : hbuf + (long long)(L - 1) * NH * n;
  const float* dz = dzbuf + ((long long)L * NH + o) * n;
  float acc[NH + 1];
#pragma unroll
  for (int i = 0; i <= NH; i++) acc[i] = 0.0f;
  for (long long s = threadIdx.x; s < n; s += 256) {
    const float d = dz[s];
    acc[NH] += d;
    if (L == 0) {
#pragma unroll
      for (int i = 0; i < NI; i++) acc[i] = __builtin_fmaf(d, hin[i * n + s], acc[i]);
    } else {
#pragma unroll
      for (int i = 0; i < NH; i++) acc[i] = __builtin_fmaf(d, hin[i * n + s], acc[i]);
    }
  }
  __shared__ float red[4][NH + 1];
#pragma unroll
  for (int i = 0; i <= NH; i++) {
    float a = acc[i];
    for (int m = 32; m > 0; m >>= 1) a += __shfl_xor(a, m);
    acc[i] = a;
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int i = 0; i <= NH; i++) red[threadIdx.x >> 6][i] = acc[i];
  }
  __syncthreads();
  const int offW[4] = {OFF_W1, OFF_W2, OFF_W3, OFF_W4};
  const int offB[4] = {OFF_B1, OFF_B2, OFF_B3, OFF_B4};
  for (int i = threadIdx.x; i <= nin; i += 256) {
    const int src = (i == nin) ? NH : i;
    const float tot = (red[0][src] + red[1][src]) + (red[2][src] + red[3][src]);
    if (i == nin) grad[offB[L] + o] = tot;
    else grad[offW[L] + o * nin + i] = tot;
  }
}

// Adam (bias-corrected step size lr_t computed on the host).
static __global__ void k_adam(float* __restrict__ P, float* __restrict__ m1, float* __restrict__ m2,
                       const float* __restrict__ G, float lr_t, float b1, float b2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NPARAM) return;
  const float g = G[i];
  const float a = b1 * m1[i] + (1.0f - b1) * g;
  const float b = b2 * m2[i] + (1.0f - b2) * g * g;
  m1[i] = a;
  m2[i] = b;
  P[i] -= lr_t * a / (sqrtf(b) + eps);
}

// FP32 parameters -> packed half-precision A operands, accumulator biases
// and the rounded weights (as float) used by the backward pass.
static __global__ void k_mlp_pack(const float* __restrict__ P, _Float16* __restrict__ packed,
                           float* __restrict__ bias, float* __restrict__ Wq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < NPARAM) {
    const bool isb = (i >= OFF_B1 && i < OFF_W2) || (i >= OFF_B2 && i < OFF_W3) ||
                     (i >= OFF_B3 && i < OFF_W4) || (i >= OFF_B4);
    Wq[i] = isb ? P[i] : h16round(P[i]);
  }
  if (i < BS_N) {
    float b;
    if (i < 64) b = P[OFF_B1 + i];
    else if (i < 128) b = P[OFF_B2 + i - 64];
    else if (i < 192) b = P[OFF_B3 + i - 128];
    else b = (i - 192 < NO) ? P[OFF_B4 + i - 192] : 0.0f;
    bias[i] = b;
  }
  if (i < PK_N) {
    float w;
    if (i < PK_L2) {  // layer 1: [rb][lane][m], k = 4g + m
      const int m = i & 3, lane = (i >> 2) & 63, rb = i >> 8;
      const int row = 16 * rb + (lane & 15), k = 4 * (lane >> 4) + m;
      w = k < NI ? P[OFF_W1 + row * NI + k] : 0.0f;
    } else {
      int j, offw, nrows;
      if (i < PK_L3) { j = i - PK_L2; offw = OFF_W2; nrows = NH; }
      else if (i < PK_L4) { j = i - PK_L3; offw = OFF_W3; nrows = NH; }
      else { j = i - PK_L4; offw = OFF_W4; nrows = NO; }
      const int m = j & 7, lane = (j >> 3) & 63, rbks = j >> 9;
      const int rb = rbks >> 1, ks = rbks & 1, gg = lane >> 4;
      const int row = 16 * rb + (lane & 15);
      const int k = 32 * ks + (m < 4 ? 4 * gg + m : 16 + 4 * gg + (m - 4));
      w = row < nrows ? P[offw + row * NH + k] : 0.0f;
    }
    packed[i] = (_Float16)w;
  }
}

}  // namespace mlp
}  // namespace echo
