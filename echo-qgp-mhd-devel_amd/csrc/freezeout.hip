// Freeze-out hypersurface finder on the GPU: the out_freeze path of
// ECHO-QGP (out_hyp, out.f90:1568-1645; hypersurface.f90) behind echo_fo_*.
//
// One slice is five launches on the grid's stream plus one device copy:
//   k_fo_field  the freeze-out field of every cell: T (freeze_type 0) or e
//               (1) from get_derived_data, into a scratch buffer, so that an
//               EoS error leaves the stored slices as they were;
//   k_fo_mask   per cell a 4-bit mask of the crossings of freeze_value along
//               tau (same cell, older and newer slice) and along x, y, eta
//               (cell and its upper neighbour on the older slice), the tests
//               of find_hypersurface_* (hypersurface.f90:172-229).  The mask
//               is stored in the reference's loop order (ix slowest, iz
//               fastest) through an LDS transpose of 32 x 32 (x, eta) tiles,
//               so the field reads (x fastest) and the mask writes (eta
//               fastest) are both coalesced;
//   k_fo_count  the element count of each block of 1024 cells in that order;
//   k_fo_scan   exclusive scan of the block counts and the total (one block);
//   k_fo_emit   every cell with crossings writes its elements (tau, x, y, eta
//               in that order) at their rank in the reference's order:
//               set_fo_{ide,vis,mhd} (354-460, 645-769, 984-1117), the sign
//               of save_hypersurface_* and the derived columns of
//               print_hypersurface_visco (832-905);
// then the older slice's primitives are replaced by the current state.
// Elements are rows of width 13 / 25 / 16 doubles (hydro / viscous / MHD).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/echo_amd.h"
#include "grid.hpp"

namespace echo {

constexpr int FO_BLOCK = 256;
constexpr int FO_SPAN = 4 * FO_BLOCK;   // cells (reference order) per scan block
constexpr int FO_TILE = 32;

template <int SYS>
struct FoWidth {
  static constexpr int W = (SysT<SYS>::MHD || SysT<SYS>::RES) ? 16 : (SysT<SYS>::VISC ? 25 : 13);
};

struct FoArgs {
  int nx, ny, nz;
  long long N;
  double fv;                 // freeze_value
  double t1, t2;             // time of the older and the newer slice
  double dx, dy, dz;         // d_val (common.f90:472-478)
  double x1, y1, z1, hx, hy, hz;   // cell centres x1 + (i - 0.5) hx, i = 1..nx
  double g1, g2, g3;         // metric at the newer slice
};

static __global__ __launch_bounds__(FO_BLOCK) void k_fo_field(GridArgs a, int ftype,
                                                             const double* __restrict__ V,
                                                             double* __restrict__ Fn,
                                                             int* __restrict__ status) {
  const long long c = (long long)blockIdx.x * FO_BLOCK + threadIdx.x;
  if (c >= a.N) return;
  double en, temp, s;
  const int err = get_derived_data(a.eos, V[KRH * a.N + c], V[KPR * a.N + c], en, temp, s);
  if (err > 0) atomicMax(status, err);
  Fn[c] = ftype == 0 ? temp : en;
}

// grid: ceil(nx/32) * ceil(nz/32) * ny blocks of 32 x 8 threads
static __global__ __launch_bounds__(FO_BLOCK) void k_fo_mask(FoArgs f, const double* __restrict__ F1,
                                                            const double* __restrict__ F2,
                                                            unsigned char* __restrict__ mask) {
  __shared__ unsigned char tile[FO_TILE][FO_TILE + 4];   // [eta][x]
  const int ntx = (f.nx + FO_TILE - 1) / FO_TILE, ntz = (f.nz + FO_TILE - 1) / FO_TILE;
  int b = blockIdx.x;
  const int bx = b % ntx;
  b /= ntx;
  const int bz = b % ntz;
  const int iy = b / ntz;
  const int ix0 = bx * FO_TILE, iz0 = bz * FO_TILE;
  const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;
  const long long sy = f.nx, sz = (long long)f.nx * f.ny;
  const double fv = f.fv;
#pragma unroll
  for (int k = 0; k < FO_TILE / 8; k++) {
    const int iz = iz0 + ly + 8 * k, ix = ix0 + lx;
    unsigned m = 0;
    if (ix < f.nx && iz < f.nz) {
      const long long c = iz * sz + iy * sy + ix;
      const double a = F1[c];
      if ((F2[c] - fv) * (fv - a) > 0.0) m |= 1u;
      if (ix < f.nx - 1 && (F1[c + 1] - fv) * (fv - a) > 0.0) m |= 2u;
      if (iy < f.ny - 1 && (F1[c + sy] - fv) * (fv - a) > 0.0) m |= 4u;
      if (iz < f.nz - 1 && (F1[c + sz] - fv) * (fv - a) > 0.0) m |= 8u;
    }
    tile[ly + 8 * k][lx] = (unsigned char)m;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < FO_TILE / 8; k++) {
    const int ix = ix0 + ly + 8 * k, iz = iz0 + lx;
    if (ix < f.nx && iz < f.nz) mask[((long long)ix * f.ny + iy) * f.nz + iz] = tile[lx][ly + 8 * k];
  }
}

__device__ __forceinline__ int fo_count4(uchar4 m) {
  return __popc(m.x) + __popc(m.y) + __popc(m.z) + __popc(m.w);
}

// inclusive scan over the 64 lanes of a wave
__device__ __forceinline__ int wave_scan(int x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x;
}

// one block per FO_SPAN cells; the mask is padded to whole blocks (zeros)
static __global__ __launch_bounds__(FO_BLOCK) void k_fo_count(const unsigned char* __restrict__ mask,
                                                             long long* __restrict__ bsum) {
  __shared__ int ws[FO_BLOCK / 64];
  const uchar4 m = reinterpret_cast<const uchar4*>(mask)[(long long)blockIdx.x * FO_BLOCK + threadIdx.x];
  const int x = wave_scan(fo_count4(m));
  if ((threadIdx.x & 63) == 63) ws[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long s = 0;
    for (int w = 0; w < FO_BLOCK / 64; w++) s += ws[w];
    bsum[blockIdx.x] = s;
  }
}

// exclusive scan of bsum[0..nb) in place, total in bsum[nb]; one block
static __global__ __launch_bounds__(1024) void k_fo_scan(long long* __restrict__ bsum, int nb) {
  __shared__ long long ws[16];
  __shared__ long long carry;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + threadIdx.x;
    const long long v = i < nb ? bsum[i] : 0;
    long long x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const long long y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    long long off = carry;
    for (int j = 0; j < w; j++) off += ws[j];
    const long long excl = off + x - v;
    if (i < nb) bsum[i] = excl;
    __syncthreads();
    if (threadIdx.x == 1023) carry = excl + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[nb] = carry;
}

// One element: the crossing along d (0 tau, 1 x, 2 y, 3 eta) from cell
// (ix, iy, iz) = c of the older slice, written to r[0 .. W).
template <int SYS, bool ZZ>
__device__ void fo_element(const GridArgs& a, const FoArgs& f, int d, int ix, int iy, int iz, long long c,
                           const double* __restrict__ F1, const double* __restrict__ F2,
                           const double* __restrict__ VB, const double* __restrict__ V,
                           double* __restrict__ r, int* __restrict__ status) {
  const long long N = f.N;
  const int jx = ix + (d == 1), jy = iy + (d == 2), jz = iz + (d == 3);
  const long long cc = c + (d == 1 ? 1 : 0) + (d == 2 ? (long long)f.nx : 0) +
                       (d == 3 ? (long long)f.nx * f.ny : 0);
  const double fa = F1[c], fb = d == 0 ? F2[c] : F1[cc];
  const double ta = f.t1, tb = d == 0 ? f.t2 : f.t1;
  const double* __restrict__ B = d == 0 ? V : VB;
  const double si = (fa - fb) >= 0. ? 1. : -1.;
  const double w = (fa - f.fv) / (fa - fb);
  auto interp = [&](int k) {
    const double A0 = VB[k * N + c];
    return A0 + w * (B[k * N + cc] - A0);
  };
  const double xa = f.x1 + ((double)(ix + 1) - 0.5) * f.hx, xb = f.x1 + ((double)(jx + 1) - 0.5) * f.hx;
  const double ya = f.y1 + ((double)(iy + 1) - 0.5) * f.hy, yb = f.y1 + ((double)(jy + 1) - 0.5) * f.hy;
  const double za = f.z1 + ((double)(iz + 1) - 0.5) * f.hz, zb = f.z1 + ((double)(jz + 1) - 0.5) * f.hz;
  const double tau = ta + w * (tb - ta);
  const double vx = interp(KVX), vy = interp(KVY), vz = interp(KVZ);
  const double rho = interp(KRH), pr = interp(KPR);
  const double betasq = vx * vx * f.g1 + vy * vy * f.g2 + vz * vz * f.g3;
  if (betasq > 1.0) atomicMax(status, ECHO_E_FREEZEOUT);
  const double dt = fabs(f.t2 - f.t1);
  r[0] = tau;
  r[1] = xa + w * (xb - xa);
  r[2] = ya + w * (yb - ya);
  r[3] = za + w * (zb - za);
  r[4] = rho;
  r[5] = vx;
  r[6] = vy;
  r[7] = vz;
  r[8] = pr;
  r[9] = si * ((double)(d == 0) * f.dx * f.dy * f.dz * tau);
  r[10] = si * ((double)(jx - ix) * dt * f.dy * f.dz * tau);
  r[11] = si * ((double)(jy - iy) * f.dx * dt * f.dz * tau);
  r[12] = si * ((double)(jz - iz) * f.dx * f.dy * dt * tau);
  if constexpr (SysT<SYS>::MHD || SysT<SYS>::RES) {
    // find_hypersurface_mhd on every MHD grid, resistive and viscous
    // resistive included (hypersurface.f90:152-166): B of the layout
    constexpr int KB = SysT<SYS>::KB;
    r[13] = interp(KB);
    r[14] = interp(KB + 1);
    r[15] = interp(KB + 2);
  } else if constexpr (SysT<SYS>::VISC) {
    double p[12];
#pragma unroll
    for (int k = 0; k < 12; k++) p[k] = 0.;
    p[KVX] = vx; p[KVY] = vy; p[KVZ] = vz;
    const double bu = interp(KPIBU), xy = interp(KPIXY), xz = interp(KPIXZ), yz = interp(KPIYZ);
    const double xx = interp(KPIXX), yy = interp(KPIYY), zz = interp(KPIZZ);
    r[13] = bu; r[14] = xy; r[15] = xz; r[16] = yz; r[17] = xx; r[18] = yy; r[19] = zz;
    p[KPIXX] = xx; p[KPIYY] = yy; p[KPIZZ] = zz; p[KPIXY] = xy; p[KPIXZ] = xz; p[KPIYZ] = yz;
    // print_hypersurface_visco: pi^{tau mu} with g33 = tau_fo^2
    // (get_derived_pi(_zz)_gcv3, viscous.f90:203-240); its EoS error flag
    // is not checked
    Metric M;
    M.g1 = 1.; M.g2 = 1.; M.g3 = tau * tau; M.g0 = -1.; M.gp = 1.; M.gm = 1.; M.t = tau;
    const PiDer q = derived_pi<ZZ>(p, M);
    double eng = 0., temp = 0., s = 0.;
    get_derived_data(a.eos, rho, pr, eng, temp, s);
    r[20] = q.pitt; r[21] = q.pitx; r[22] = q.pity; r[23] = q.pitz;
    r[24] = 1.0 / (temp * temp * (eng + pr));
  }
}

template <int SYS, bool ZZ>
__global__ __launch_bounds__(FO_BLOCK) void k_fo_emit(GridArgs a, FoArgs f, const unsigned char* __restrict__ mask,
                                                      const long long* __restrict__ bsum,
                                                      const double* __restrict__ F1,
                                                      const double* __restrict__ F2,
                                                      const double* __restrict__ VB,
                                                      const double* __restrict__ V,
                                                      double* __restrict__ rec, int* __restrict__ status) {
  constexpr int W = FoWidth<SYS>::W;
  __shared__ int ws[FO_BLOCK / 64];
  const uchar4 m4 = reinterpret_cast<const uchar4*>(mask)[(long long)blockIdx.x * FO_BLOCK + threadIdx.x];
  const int n = fo_count4(m4);
  const int x = wave_scan(n);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 63) ws[w] = x;
  __syncthreads();
  if (n == 0) return;
  long long pos = bsum[blockIdx.x] + x - n;
  for (int j = 0; j < w; j++) pos += ws[j];
  const unsigned char ms[4] = {m4.x, m4.y, m4.z, m4.w};
  const long long r0 = (long long)blockIdx.x * FO_SPAN + 4 * threadIdx.x;
  const long long nyz = (long long)f.ny * f.nz;
  for (int j = 0; j < 4; j++) {
    const unsigned m = ms[j];
    if (!m) continue;
    const long long rr = r0 + j;
    const int ix = (int)(rr / nyz);
    const int iy = (int)((rr / f.nz) % f.ny);
    const int iz = (int)(rr % f.nz);
    const long long c = ((long long)iz * f.ny + iy) * f.nx + ix;
    for (int d = 0; d < 4; d++)
      if (m >> d & 1u) {
        fo_element<SYS, ZZ>(a, f, d, ix, iy, iz, c, F1, F2, VB, V, rec + pos * W, status);
        pos++;
      }
  }
}

// Fortran ES14.7
static void es14_7(char* b, size_t nb, double x) {
  char t[40];
  snprintf(t, sizeof t, "%.7E", x);
  char* e = strchr(t, 'E');
  if (e && strlen(e + 2) > 2) memmove(e, e + 1, strlen(e + 1) + 1);   // +ddd: no E
  snprintf(b, nb, "%14s", t);
}

// list-directed output of a real(8) as the flang runtime writes it: the
// shortest digits that read back to x, F form for 0.095 <= |x| < 1e15
// (no leading zero), E form with at least two exponent digits otherwise
static std::string list_real(double x) {
  if (x == 0.) return "0.";
  char t[40];
  for (int p = 1; p <= 17; p++) {
    snprintf(t, sizeof t, "%.*e", p - 1, x);
    if (strtod(t, nullptr) == x) break;
  }
  std::string dig;
  const char* q = t + (t[0] == '-');
  for (; *q && *q != 'e'; q++)
    if (*q != '.') dig += *q;
  const int ex = atoi(q + 1);
  const int nd = (int)dig.size();
  std::string o = x < 0 ? "-" : "";
  const bool f = (ex >= -1 && ex < 15) || (ex == -2 && dig[0] == '9' && nd > 1 && dig[1] >= '5');
  if (f) {
    if (ex >= 0) {
      for (int i = 0; i <= ex; i++) o += i < nd ? dig[i] : '0';
      o += '.';
      for (int i = ex + 1; i < nd; i++) o += dig[i];
    } else {
      o += '.';
      for (int i = 0; i < -ex - 1; i++) o += '0';
      o += dig;
    }
  } else {
    o += dig[0];
    o += '.';
    o += dig.substr(1);
    char e[16];
    snprintf(e, sizeof e, "E%c%02d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
    o += e;
  }
  return o;
}

}  // namespace echo

using namespace echo;

extern "C" {

int echo_fo_width(const echo_grid* g) {
  if (!g) return 0;
  return g->cfg.mhd ? 16 : (g->cfg.viscosity ? 25 : 13);
}

int echo_fo_setup(echo_grid* g, int32_t freeze_type, double freeze_value, double* hdr) {
  if (!g) return fail(ECHO_E_ARG, "echo_fo_setup: null grid");
  if (freeze_type != 0 && freeze_type != 1)
    return fail(ECHO_E_ARG, "freeze_type must be 0 (temperature) or 1 (energy density), got %d", freeze_type);
  if (!isfinite(freeze_value)) return fail(ECHO_E_ARG, "freeze_value must be finite");
  HIPCHECK(hipSetDevice(g->dev));
  const long long N = g->N;
  const long long nb = (N + FO_SPAN - 1) / FO_SPAN;
  if (!g->fo_f[0]) {
    void* p[7] = {};
    const size_t sz[7] = {sizeof(double) * N, sizeof(double) * N, sizeof(double) * N,
                          sizeof(double) * N * g->nv, (size_t)nb * FO_SPAN,
                          sizeof(long long) * (nb + 1), sizeof(int)};
    for (int i = 0; i < 7; i++)
      if (hipMalloc(&p[i], sz[i]) != hipSuccess) {
        for (int j = 0; j < i; j++) hipFree(p[j]);
        return fail(ECHO_E_NOMEM, "echo_fo_setup: device allocation of %zu bytes failed", sz[i]);
      }
    g->fo_f[0] = (double*)p[0];
    g->fo_f[1] = (double*)p[1];
    g->fo_f[2] = (double*)p[2];
    g->fo_vb = (double*)p[3];
    g->fo_mask = (unsigned char*)p[4];
    g->fo_bsum = (long long*)p[5];
    g->fo_status = (int*)p[6];
    HIPCHECK(hipMemset(g->fo_mask, 0, (size_t)nb * FO_SPAN));
  }
  g->fo_on = true;
  g->fo_type = freeze_type;
  g->fo_value = freeze_value;
  g->fo_t1 = g->fo_t2 = 0.;
  g->fo_n = 0;
  g->fo_searched = false;
  // get_hysufile_ready (hypersurface.f90:1155-1180); the reference hands an
  // unset density to the EoS, which the tabulated and p = e/3 EoS ignore
  const Eos E = host_eos(g);
  double temp_f = 0., eng_f = 0., prex = 0., d1, d2, s;
  int err = 0;
  if (freeze_type == 0) {
    temp_f = freeze_value;
    prex = pressure_from_temperature(E, temp_f);
    err = eos_energy(E, 1., prex, eng_f);
  } else {
    eng_f = freeze_value;
    err = eos_pressure(E, 1., eng_f, prex, d1, d2);
    if (err <= 0) err = get_derived_data(E, 1., prex, eng_f, temp_f, s);
  }
  g->fo_hdr[0] = temp_f;
  g->fo_hdr[1] = eng_f;
  if (hdr) { hdr[0] = temp_f; hdr[1] = eng_f; }
  if (err > 0) return fail(err, "echo_fo_setup: EoS error %d at the freeze-out value", err);
  return 0;
}

int echo_fo_slice(echo_grid* g, int64_t* n_elements) {
  if (!g || !n_elements) return fail(ECHO_E_ARG, "echo_fo_slice: null argument");
  *n_elements = 0;
  if (!g->fo_on) return fail(ECHO_E_ARG, "echo_fo_slice: call echo_fo_setup first");
  HIPCHECK(hipSetDevice(g->dev));
  hipStream_t st = g->stream;
  const long long N = g->N;
  const int nb = (int)((N + FO_SPAN - 1) / FO_SPAN);
  double t = 0.;
  int status = 0;
  HIPCHECK(hipMemcpyAsync(&t, &g->S->t, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipMemsetAsync(g->fo_status, 0, sizeof(int), st));
  k_fo_field<<<nblocks(N, FO_BLOCK), FO_BLOCK, 0, st>>>(g->ga, g->fo_type, g->V, g->fo_f[2], g->fo_status);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(&status, g->fo_status, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  if (status > 0) return fail(status, "freeze-out field: EoS error %d", status);
  // out_hyp: field(1) = field(2), field(2) = the new field
  const bool first = g->fo_t2 == 0.;
  if (first) g->fo_t2 = t;
  double* old = g->fo_f[0];
  g->fo_f[0] = g->fo_f[1];
  g->fo_f[1] = g->fo_f[2];
  g->fo_f[2] = old;
  g->fo_n = 0;
  g->fo_searched = !first;
  const size_t vbytes = sizeof(double) * N * g->nv;
  if (first) {
    HIPCHECK(hipMemcpyAsync(g->fo_vb, g->V, vbytes, hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipStreamSynchronize(st));
    return 0;
  }
  g->fo_t1 = g->fo_t2;
  g->fo_t2 = t;
  const echo_config& c = g->cfg;
  FoArgs f;
  f.nx = c.nx; f.ny = c.ny; f.nz = c.nz; f.N = N;
  f.fv = g->fo_value;
  f.t1 = g->fo_t1; f.t2 = g->fo_t2;
  f.hx = (c.x2 - c.x1) / c.nx; f.hy = (c.y2 - c.y1) / c.ny; f.hz = (c.z2 - c.z1) / c.nz;
  f.dx = f.hx; f.dy = f.hy; f.dz = c.nz > 1 ? f.hz : 1.0;
  f.x1 = c.x1; f.y1 = c.y1; f.z1 = c.z1;
  const Metric M = make_metric(c.coordinates, t);
  f.g1 = M.g1; f.g2 = M.g2; f.g3 = M.g3;
  const long long ntile = (long long)((c.nx + FO_TILE - 1) / FO_TILE) * ((c.nz + FO_TILE - 1) / FO_TILE) * c.ny;
  if (ntile > 0x7fffffffLL) return fail(ECHO_E_ARG, "grid too large for the freeze-out finder");
  k_fo_mask<<<(unsigned)ntile, FO_BLOCK, 0, st>>>(f, g->fo_f[0], g->fo_f[1], g->fo_mask);
  k_fo_count<<<nb, FO_BLOCK, 0, st>>>(g->fo_mask, g->fo_bsum);
  k_fo_scan<<<1, 1024, 0, st>>>(g->fo_bsum, nb);
  HIPCHECK(hipGetLastError());
  long long total = 0;
  HIPCHECK(hipMemcpyAsync(&total, g->fo_bsum + nb, sizeof(long long), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  const int W = echo_fo_width(g);
  if (total > g->fo_cap) {
    const long long cap = total + total / 2 + 1024;
    hipFree(g->fo_rec);
    g->fo_rec = nullptr;
    g->fo_cap = 0;
    if (hipMalloc(&g->fo_rec, sizeof(double) * W * cap) != hipSuccess)
      return fail(ECHO_E_NOMEM, "freeze-out elements: allocation of %lld rows failed", cap);
    g->fo_cap = cap;
  }
  if (total > 0) {
    dispatch_sys(c, [&](auto sys, auto zz) {
      k_fo_emit<decltype(sys)::value, decltype(zz)::value><<<nb, FO_BLOCK, 0, st>>>(
          g->ga, f, g->fo_mask, g->fo_bsum, g->fo_f[0], g->fo_f[1], g->fo_vb, g->V, g->fo_rec, g->fo_status);
    });
    HIPCHECK(hipGetLastError());
  }
  // v_backup = primitives: the current state becomes the older slice
  HIPCHECK(hipMemcpyAsync(g->fo_vb, g->V, vbytes, hipMemcpyDeviceToDevice, st));
  HIPCHECK(hipMemcpyAsync(&status, g->fo_status, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  g->fo_n = total;
  *n_elements = total;
  if (status == ECHO_E_FREEZEOUT)
    return fail(ECHO_E_FREEZEOUT, "freeze-out element with v^2 > 1 (the reference stops there)");
  return 0;
}

int echo_fo_elements(echo_grid* g, double* out, int64_t n) {
  if (!g || (!out && n > 0)) return fail(ECHO_E_ARG, "echo_fo_elements: null argument");
  if (n < 0 || n > g->fo_n) return fail(ECHO_E_ARG, "echo_fo_elements: %lld rows asked, %lld stored",
                                        (long long)n, g->fo_n);
  if (n == 0) return 0;
  HIPCHECK(hipSetDevice(g->dev));
  HIPCHECK(hipMemcpyAsync(out, g->fo_rec, sizeof(double) * echo_fo_width(g) * n, hipMemcpyDeviceToHost,
                          g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  return 0;
}

int echo_fo_write(echo_grid* g, const char* path, int32_t header) {
  if (!g || !path) return fail(ECHO_E_ARG, "echo_fo_write: null argument");
  if (!g->fo_on) return fail(ECHO_E_ARG, "echo_fo_write: call echo_fo_setup first");
  if (!header && !g->fo_searched) return 0;   // the first slice is only stored
  const int W = echo_fo_width(g);
  std::vector<double> rows;
  if (!header && g->fo_n > 0) {
    rows.resize((size_t)g->fo_n * W);
    int e = echo_fo_elements(g, rows.data(), g->fo_n);
    if (e) return e;
  }
  FILE* fp = fopen(path, header ? "w" : "a");
  if (!fp) return fail(ECHO_E_ARG, "echo_fo_write: cannot open %s", path);
  if (header) {
    fprintf(fp, " %d %s %s\n", g->fo_type, list_real(g->fo_hdr[0]).c_str(), list_real(g->fo_hdr[1]).c_str());
  } else {
    if (g->fo_n > 999999) fputs("******\n", fp);
    else fprintf(fp, "%6lld\n", g->fo_n);
    std::string line;
    char v[40];
    for (long long i = 0; i < g->fo_n; i++) {
      line.clear();
      for (int j = 0; j < W; j++) {
        es14_7(v, sizeof v, rows[(size_t)i * W + j]);
        line += v;
        line += j < W - 1 ? "  " : "\n";
      }
      fputs(line.c_str(), fp);
    }
  }
  const bool bad = ferror(fp) != 0;
  if (fclose(fp) != 0 || bad) return fail(ECHO_E_ARG, "echo_fo_write: write to %s failed", path);
  return 0;
}

}  // extern "C"
