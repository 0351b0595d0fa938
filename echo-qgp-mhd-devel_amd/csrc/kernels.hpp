// Kernels of one explicit RK stage of the ECHO-QGP grid update on gfx950.
//
// Data layout in HBM (FP64, variable-major SoA, x fastest; the reference's
// Fortran layout of v/u, common.f90:203-214):
//   V            nv * N       primitives (updated in place by k_update)
//   U[0], U[1]   nv * N each  conserved u^n and the stage value
//   F[d]         nv * Nf[d]   HLL flux on the interfaces normal to d
//   L[d]         4 * Nf[d]    face sums wdl+wdr of (gamma, gamma v^i), viscous
//   VS[0], VS[1] 3 * N        v^i at the start of the current / previous step
//                             (vnewstep / vold of echo.f90:192-241), viscous
// Interfaces normal to x are indexed i + (nx+1)*(y + ny*z), i = 0..nx, the
// interface i lying between cells i-1 and i (0-based); likewise for y and z
// with the interface index in its own dimension and x always fastest, so a
// wavefront of consecutive threads reads consecutive x.
#pragma once
#include "source.hpp"

namespace echo {

struct DevScalars {
  double t, dt, timeold, timeinterval, g_cov_3_old, tout, glm_ch, glm_ch_old;
  double smooth_pressure, dtout, tmax;
  double temp_end;         // check_temp threshold (<= 0: no check)
  double aflux[3];         // max(ap, am) at the last interface of each sweep
  double check_max;        // max temperature (or energy density) of the last check
  long long nrepair;       // cells repaired by k_fixup since set-up
  long long nfallback;     // states the learned inversion in the step left to the plain solver
  int iter, run_crashed, errcode, nfail;
  int vs_parity;
  int freeze_type;         // 0: temperature, 1: energy density
  int temp_check;          // 1 once the check has stopped the run
  int crash_kind;          // what crashed the run: CRASH_* below
  int crash_cell;          // first cell (x fastest) whose repair failed
  int pad;
};

enum { CRASH_NONE = 0, CRASH_REPAIR = 1, CRASH_FLUX = 2, CRASH_SOURCE = 3 };

// Scratch of the con2prim repair (k_update -> k_fixup), valid within one
// stage; an ensemble shares one set sized for its largest event.
//   Pold     N        the pressure each cell held before the stage
//   Upre     nv * N   only with the viscous smoothing: u of the smoothed cells
//                     before the smoothing
//   failbits N / 32   one bit per cell whose con2prim failed (cleared by k_fixup)
//   list     N        the failed cells in increasing order (k_fixup)
//   level    N        dependency level of each failed cell (k_fixup)
struct FixBufs {
  double* Pold;
  double* Upre;
  unsigned* failbits;
  int* list;
  int* level;
};

struct GridArgs {
  int nx, ny, nz, coordinates, bc;
  long long N;
  double ddx, ddy, ddz, maxspeed, cfl, maxdt;
  Eos eos;
  ViscParams vp;
  int enable_smooth;
  int algo_solver;      // MHD con2prim algorithm (4, 5 or 6)
  int divclean;         // Dedner GLM on
  double glm_alpha;     // GLM damping parameter
  int res;              // resistive MHD (E evolved): light-speed CFL, glm_ch = 1
  double sig, sch;      // electric and chiral conductivity, fm^-1
  int ext_c2p;          // ideal MHD: k_update leaves con2prim to the learned inversion
  int upd_tz;           // z rows per block of the cell updates (update_cell): 4 or 1
  DevScalars* S;
};

// IMEX-SSP3(4,3,3) scratch (resistive systems), nv components each, in
// units of a rate (u = u0 - dt * rate).  Only the three E components carry
// a stiff part, so the other components need one slot between stages:
//   P4   E: the explicit start of K4 as it accumulates (b ds1, + du2/4 + h ds2);
//        other components: du2 / 4 after stage 2, then R3 = du2/4 + du3/4
//        (the rate of K4's start) after stage 3;
//   ACC  E: the final combination as it accumulates (ds2/6 + du2/6, + du3/6
//        + ds3/6); the other components take it from P4 (2/3 R3).
struct ImexBufs {
  double* P4;
  double* ACC;
};

// What the con2prim repair inside an implicit IMEX stage reads as a
// neighbour's explicit start U*: the solve kernels store U* only for the
// three E components (and the repair stores a repaired cell's whole u); the
// other components follow from u0 and the stored rate with the arithmetic
// the update used:
//   phase 1 (K1) U* = u0 (E too), phase 2 (K2) u0,
//   phase 3 (K3) fma(-dt, 4 P4, u0), phase 4 (K4) fma(-dt, P4, u0).
struct ImexView {
  const double* U0;
  const double* P4;
  int phase;              // 0: not inside an implicit stage
};

// 128 threads (two waves) per flux block: with 24 KB of LDS per block, six
// blocks share a CU at three waves per SIMD, and the smaller blocks even out
// the tail of a sweep (about 1% faster sweeps than 256)
constexpr int FLUX_BLOCK = 128;
constexpr int UPDATE_BLOCK = 256;
constexpr int FIX_BLOCK = 256;
constexpr int NXCD = 8;   // MI355X: 8 XCDs, each with its own L2

// XCD-aware block order.  Workgroups are dispatched round-robin over the
// XCDs (block b runs on XCD b % 8), so neighbouring blocks, which share
// stencil rows, would sit in different L2 caches.  Launch grids are padded
// to a multiple of NXCD and block b takes the logical tile
// (b % NXCD) * (nb / NXCD) + b / NXCD: each XCD then walks one contiguous
// eighth of the index space.  A bijection on [0, nb).
__device__ __forceinline__ long long xcd_block(unsigned b, unsigned nb) {
  const unsigned per = nb / NXCD;
  return (long long)(b % NXCD) * per + b / NXCD;
}

// Cell of a cell-update thread (k_update, k_update_imex).  With upd_tz = 4 a
// block is a tile of 64 x-consecutive cells by 4 z-consecutive rows (one
// wave per row), tiles ordered x chunk fastest, then y, then z.  A cell
// reads the z faces below and above it, and the face between two rows of a
// tile is read by both rows within the block: the face arrays of the z
// sweep are fetched about once, not once per side (with the cells in
// storage order the row above comes nx * ny cells later, long after the
// face has left the XCD's L2).  MHD grids with at least 64 cells in x and
// 4 in z use the tiles; with upd_tz = 1 the cells go in storage order,
// UPDATE_BLOCK per block.
// Returns false for the padding threads of the launch.
constexpr int UPD_TX = 64, UPD_TZ = 4;
static_assert(UPDATE_BLOCK == UPD_TX * UPD_TZ && UPD_TX == 64, "a tile row is one wave of 64 lanes");
__device__ __forceinline__ bool update_cell(const GridArgs& a, int& x, int& y, int& z, long long& c) {
  const long long b = xcd_block(blockIdx.x, gridDim.x);
  if (a.upd_tz == 1) {
    c = b * UPDATE_BLOCK + threadIdx.x;
    if (c >= a.N) return false;
    x = (int)(c % a.nx);
    const long long r = c / a.nx;
    y = (int)(r % a.ny);
    z = (int)(r / a.ny);
    return true;
  }
  const int nxc = (a.nx + UPD_TX - 1) / UPD_TX;
  const long long r = b / nxc;
  x = (int)(b % nxc) * UPD_TX + (int)(threadIdx.x & (UPD_TX - 1));
  y = (int)(r % a.ny);
  const long long zz = (r / a.ny) * UPD_TZ + (threadIdx.x >> 6);
  if (x >= a.nx || zz >= a.nz) return false;
  z = (int)zz;
  c = x + (long long)a.nx * (y + (long long)a.ny * z);
  return true;
}

// cell index along a direction of length m for a stencil position c0
// (0-based, may lie in the ghost range -3..m+2): boundary conditions of
// evolve_bc{x,y,z}, evolve.f90:872-1127 (0 periodic, 2 zeroth-order outflow)
__device__ __forceinline__ int bc_map(int c0, int m, int bc) {
  if (bc == 0) return c0 < 0 ? c0 + m : (c0 >= m ? c0 - m : c0);
  return c0 < 0 ? 0 : (c0 >= m ? m - 1 : c0);
}

// ---------------------------------------------------------------------------
// k_flux: one thread per interface normal to DIR.  Reconstructs the nv
// primitives (WENO5/PPM4/TVD2/...) from the six cells around the interface,
// applies the reference's repair of unphysical states (evolve.f90:334-369),
// evaluates the HLL flux (work.f90:195-229 with system_flux of
// system.f90:1316-1583) and writes it, plus a per-block max of the fast
// speeds for the CFL condition (evolve.f90:723-764).
// ---------------------------------------------------------------------------
//
// Tile shape (TY): with TY = 1 a block takes FLUX_BLOCK consecutive faces in
// the face order (x fastest).  For the y and z sweeps TY > 1 makes the block
// a tile of FLUX_BLOCK/TY faces along x by TY faces along the sweep, so the
// six-cell stencils of the block's faces overlap in TY + 5 cell rows instead
// of 6 * TY; the autotuner picks TY per sweep and grid.
template <int DIR>
__host__ __device__ __forceinline__ long long flux_tiles(int nx, int ny, int nz, int ty) {
  const long long nf = DIR == 1 ? (long long)(nx + 1) * ny * nz
                     : (DIR == 2 ? (long long)nx * (ny + 1) * nz : (long long)nx * ny * (nz + 1));
  if (ty == 1 || DIR == 1) return (nf + FLUX_BLOCK - 1) / FLUX_BLOCK;
  const int tx = FLUX_BLOCK / ty;
  const long long ntx = (nx + tx - 1) / tx;
  const long long nti = ((DIR == 2 ? ny : nz) + 1 + ty - 1) / ty;
  return ntx * nti * (DIR == 2 ? nz : ny);
}

// Launch bounds (FLUX_BLOCK, 3): at most 168 VGPRs, so three waves share
// each SIMD.  Unbounded, the viscous WENO5 kernel takes 320 VGPRs (one wave
// per SIMD) and its FP64 chains and stencil loads stall with nothing to
// switch to.  The reconstructed states are what does not fit: 2 x NV doubles
// per lane beside the stencils in flight and the side states.  They wait in
// LDS instead, sw[2][NV][FLUX_BLOCK] (24 KB per block for NV = 12: 144 KB
// for the six blocks of a CU at three waves per SIMD).  In viscous runs the pi^{mu nu} components go there as
// soon as they are reconstructed; rho, v and p stay in registers for the
// repair.  Each lane touches only its own column of sw, so no barrier is
// needed.  Every variant then fits without scratch spills; before this, the
// viscous WENO5 kernel spilled ~48 VGPRs per lane to scratch, which lives in
// HBM (about 1.7 GB of extra traffic per sweep on 192 x 192 x 128).  The
// viscous resistive system (NV = 19, 39 KB of LDS per block) fits four
// blocks per CU, two waves per SIMD, and is bounded for that.
template <int SYS>
constexpr int flux_waves() { return SysT<SYS>::NV > 12 ? 2 : 3; }

template <int DIR, int REC, int SYS, bool ZZ, int TY = 1>
__global__ __launch_bounds__(FLUX_BLOCK, flux_waves<SYS>()) void k_flux(GridArgs a, const double* __restrict__ V,
                                                     double* __restrict__ F, double* __restrict__ L,
                                                     double* __restrict__ partial) {
  constexpr bool VISC = SysT<SYS>::VISC;
  constexpr int NV = SysT<SYS>::NV;
  static_assert(TY == 1 || DIR != 1, "x sweeps use TY = 1");
  if (a.S->temp_check) return;   // the run has stopped (check_temp)
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int m = DIR == 1 ? nx : (DIR == 2 ? ny : nz);
  const long long nf = DIR == 1 ? (long long)(nx + 1) * ny * nz
                     : (DIR == 2 ? (long long)nx * (ny + 1) * nz : (long long)nx * ny * (nz + 1));
  const long long tile = xcd_block(blockIdx.x, gridDim.x);
  long long f;
  bool inside;
  if constexpr (TY == 1) {
    f = tile * FLUX_BLOCK + threadIdx.x;
    inside = f < nf;
  } else {
    constexpr int TX = FLUX_BLOCK / TY;
    const long long ntx = (nx + TX - 1) / TX;
    const long long nti = (m + 1 + TY - 1) / TY;
    const long long r = tile / ntx;
    const int x = (int)(tile % ntx) * TX + (int)threadIdx.x % TX;
    const int i = (int)(r % nti) * TY + (int)threadIdx.x / TX;
    const long long rest = r / nti;   // z (y sweep) or y (z sweep)
    inside = x < nx && i <= m && rest < (DIR == 2 ? nz : ny);
    f = DIR == 2 ? x + (long long)nx * (i + (long long)(ny + 1) * rest)
                 : x + (long long)nx * (rest + (long long)ny * i);
  }
  double amax = 0.;
  if (inside) {
    int x, y, z, i;
    long long base, stride;
    if constexpr (DIR == 1) {
      i = (int)(f % (nx + 1)); long long r = f / (nx + 1); y = (int)(r % ny); z = (int)(r / ny);
      base = ((long long)z * ny + y) * nx; stride = 1;
    } else if constexpr (DIR == 2) {
      x = (int)(f % nx); long long r = f / nx; i = (int)(r % (ny + 1)); z = (int)(r / (ny + 1));
      base = (long long)z * ny * nx + x; stride = nx;
    } else {
      x = (int)(f % nx); long long r = f / nx; y = (int)(r % ny); i = (int)(r / ny);
      base = (long long)y * nx + x; stride = (long long)nx * ny;
    }
    (void)x; (void)y; (void)z;
    // interface i lies between cells i-1 and i (0-based); stencil cells i-3 .. i+2
    long long off[6];
#pragma unroll
    for (int j = 0; j < 6; j++) off[j] = base + stride * bc_map(i - 3 + j, m, a.bc);

    const DevScalars* S = a.S;
    const Metric M = make_metric(a.coordinates, S->t);
    if constexpr (VISC) {
      // Lorentz-velocity reconstruction feeding deriv, evolve.f90:324-331,
      // 371-374; done first so that its 42 temporaries are dead before the
      // nv reconstructions below
      double wd[4][6];
#pragma unroll
      for (int j = 0; j < 6; j++) {
        const double vx = V[(long long)KVX * a.N + off[j]];
        const double vy = V[(long long)KVY * a.N + off[j]];
        const double vz = V[(long long)KVZ * a.N + off[j]];
        const double gl = rsqrt_nr(1. - vx * vx - vy * vy - M.g3 * (vz * vz));
        wd[0][j] = gl; wd[1][j] = gl * vx; wd[2][j] = gl * vy; wd[3][j] = gl * vz;
      }
      const long long nfl = nf;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        double l, r;
        rec_face<REC>(wd[q], l, r);
        // outputs are streamed: non-temporal stores keep them from
        // evicting the stencil rows that neighbouring blocks still read
        __builtin_nontemporal_store(l + r, &L[q * nfl + f]);
      }
    }
    // c0: the cell left of the interface, used by the repair (KVX..KPR).
    // Components k >= KREG of wl, wr go straight to sw.
    __shared__ double sw[2][NV][FLUX_BLOCK];
    constexpr int KREG = VISC ? KPR + 1 : NV;
    double wl[NV], wr[NV], c0[5];
    double dEface = 0.;   // resistive: E^d(right cell) - E^d(left cell), for div E
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s[6];
      const double* Vk = V + (long long)k * a.N;
#pragma unroll
      for (int j = 0; j < 6; j++) s[j] = Vk[off[j]];
      double l, r;
      rec_face<REC>(s, l, r);
      if (k < KREG) { wl[k] = l; wr[k] = r; } else { sw[0][k][threadIdx.x] = l; sw[1][k][threadIdx.x] = r; }
      if (k <= KPR) c0[k] = s[2];
      if constexpr (SysT<SYS>::RES) {
        if (k == SysT<SYS>::KE + DIR - 1) dEface = s[3] - s[2];
      }
    }
    // repair of reconstructed states with the cell left of the interface
    // (evolve.f90:334-369; the reference uses w0(ix) for both sides)
    {
      if (wr[KPR] <= 0) wr[KPR] = c0[KPR];
      double v2 = wr[KVX] * wr[KVX] * M.g1 + wr[KVY] * wr[KVY] * M.g2 + wr[KVZ] * wr[KVZ] * M.g3;
      if (v2 >= 1.) {
        wr[KVX] = c0[KVX]; wr[KVY] = c0[KVY]; wr[KVZ] = c0[KVZ];
        v2 = wr[KVX] * wr[KVX] * M.g1 + wr[KVY] * wr[KVY] * M.g2 + wr[KVZ] * wr[KVZ] * M.g3;
        if (v2 >= 1) {
          double s = sqrt(v2);
#pragma unroll
          for (int k = KVX; k <= KVZ; k++) { c0[k] = c0[k] * a.maxspeed / s; wr[k] = c0[k]; }
        }
      }
      if (wl[KPR] <= 0) wl[KPR] = c0[KPR];
      v2 = wl[KVX] * wl[KVX] * M.g1 + wl[KVY] * wl[KVY] * M.g2 + wl[KVZ] * wl[KVZ] * M.g3;
      if (v2 >= 1.) {
        wl[KVX] = c0[KVX]; wl[KVY] = c0[KVY]; wl[KVZ] = c0[KVZ];
        v2 = wl[KVX] * wl[KVX] * M.g1 + wl[KVY] * wl[KVY] * M.g2 + wl[KVZ] * wl[KVZ] * M.g3;
        if (v2 >= 1) {
          double s = sqrt(v2);
#pragma unroll
          for (int k = KVX; k <= KVZ; k++) { c0[k] = c0[k] * a.maxspeed / s; wl[k] = c0[k]; }
        }
      }
    }
    if constexpr (SysT<SYS>::GLM) {
      // work_glm{x,y,z}, work.f90:67-139: Dedner interface values of psi and
      // the normal B.  The reference sets wr(psi) = wl(psi) before taking
      // their difference, so the B update reduces to the average.  In the
      // viscous resistive system these components wait in LDS.
      constexpr int KB = SysT<SYS>::KB + DIR - 1, KG = SysT<SYS>::KG;
      const double ch = S->glm_ch;
      double gl, gr, bl, br;
      if constexpr (KG < KREG) { gl = wl[KG]; gr = wr[KG]; bl = wl[KB]; br = wr[KB]; }
      else {
        gl = sw[0][KG][threadIdx.x]; gr = sw[1][KG][threadIdx.x];
        bl = sw[0][KB][threadIdx.x]; br = sw[1][KB][threadIdx.x];
      }
      gl = 0.5 * (gl + gr) - (br - bl) * ch / 2;
      gr = gl;
      bl = 0.5 * (bl + br) - (gr - gl) / (2 * ch);
      br = bl;
      if constexpr (KG < KREG) { wl[KG] = gl; wr[KG] = gr; wl[KB] = bl; wr[KB] = br; }
      else {
        sw[0][KG][threadIdx.x] = gl; sw[1][KG][threadIdx.x] = gr;
        sw[0][KB][threadIdx.x] = bl; sw[1][KB][threadIdx.x] = br;
      }
    }
    // Park wr in LDS while the left side state is formed, then wl while the
    // right one is.  side_state completes the pi^{tau mu} components of its
    // input (derived_pi), so each side is stored after its own side_state.
    // The empty asm statements with a memory clobber keep the compiler from
    // forwarding the stored values in registers across the side-state
    // computations, which is what would spill.
#pragma unroll
    for (int k = 0; k < KREG; k++) sw[1][k][threadIdx.x] = wr[k];
#pragma unroll
    for (int k = KREG; k < NV; k++) wl[k] = sw[0][k][threadIdx.x];
    SideState sl, sr;
    int e = side_state<DIR, SYS, ZZ>(a.eos, M, wl, sl);
#pragma unroll
    for (int k = 0; k < NV; k++) sw[0][k][threadIdx.x] = wl[k];
    asm volatile("" ::: "memory");
    {
      double wv[NV];
#pragma unroll
      for (int k = 0; k < NV; k++) wv[k] = sw[1][k][threadIdx.x];
      if (e == 0) e = side_state<DIR, SYS, ZZ>(a.eos, M, wv, sr);
#pragma unroll
      for (int k = 0; k < NV; k++) sw[1][k][threadIdx.x] = wv[k];
    }
    asm volatile("" ::: "memory");
    if (e != 0) {
      // the reference stops the run (run_crashed) on a flux error
      atomicMax(&a.S->errcode, e);
      atomicExch(&a.S->run_crashed, 1);
      atomicCAS(&a.S->crash_kind, CRASH_NONE, CRASH_FLUX);
#pragma unroll
      for (int k = 0; k < NV; k++) F[(long long)k * nf + f] = 0.;
    } else {
      // HLL, work.f90:195-229
      double vfl0, vfl1, vfr0, vfr1;
      side_speeds<DIR>(M, sl, vfl0, vfl1);
      side_speeds<DIR>(M, sr, vfr0, vfr1);
      double ap = dmax(dmax(0., vfl0), vfr0);
      double am = dmax(dmax(0., -vfl1), -vfr1);
      amax = dmax(ap, am);
      // aflux(dir) keeps the last interface of the sweep (work.f90:217,310,401)
      if (f == nf - 1) a.S->aflux[DIR - 1] = amax;
      double a1 = (ap == 0.0 && am == 0.0) ? 1. : rcp_nr(ap + am);
      // resistive: B, E and psi with the speed of light (coordinate speed
      // 1 / sqrt(g_dir))
      const double clight = DIR == 3 ? M.gm : 1.;
#pragma unroll
      for (int k = 0; k < NV; k++) {
        double ql[NV], qr[NV];
#pragma unroll
        for (int j = 0; j < NV; j++) { ql[j] = sw[0][j][threadIdx.x]; qr[j] = sw[1][j][threadIdx.x]; }
        double ulk = u_comp<SYS>(k, M, sl, ql) * M.gp;
        double urk = u_comp<SYS>(k, M, sr, qr) * M.gp;
        double flk = f_comp<DIR, SYS>(k, M, sl, ql) * M.gp;
        double frk = f_comp<DIR, SYS>(k, M, sr, qr) * M.gp;
        double fk;
        if (SysT<SYS>::RES && k >= SysT<SYS>::KB) fk = 0.5 * (flk + frk - clight * (urk - ulk));
        else fk = a1 * (ap * flk + am * frk - ap * am * (urk - ulk));
        __builtin_nontemporal_store(fk, &F[(long long)k * nf + f]);
      }
      if constexpr (SysT<SYS>::RES) __builtin_nontemporal_store(dEface, &F[(long long)NV * nf + f]);
    }
  }
  // block max of the interface speeds -> partial[blockIdx.x]
  __shared__ double red[FLUX_BLOCK / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) amax = dmax(amax, __shfl_xor(amax, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = amax;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
#pragma unroll
    for (int w = 1; w < FLUX_BLOCK / 64; w++) r = dmax(r, red[w]);
    partial[blockIdx.x] = r;
  }
}

// ---------------------------------------------------------------------------
// k_cfl: reduces the per-block speed maxima of the three sweeps and applies
// evolve_dt (evolve.f90:723-764) and the clamps of evolve.f90:179-188.
// One block.  partial holds nb0 + nb1 + nb2 values (the block counts of the
// three sweeps, in this order); only the sweeps of active dimensions
// (n > 1) wrote theirs.
// ---------------------------------------------------------------------------
static __global__ __launch_bounds__(1024) void k_cfl(GridArgs a, const double* __restrict__ partial,
                                              int nb0, int nb1, int nb2) {
  if (a.S->temp_check) return;
  __shared__ double red[3][16];
  double m[3] = {0., 0., 0.};
  const int nb[3] = {nb0, nb1, nb2};
  const bool act[3] = {a.nx > 1, a.ny > 1, a.nz > 1};
  int off = 0;
  for (int d = 0; d < 3; d++) {
    if (act[d])
      for (int i = threadIdx.x; i < nb[d]; i += blockDim.x) m[d] = dmax(m[d], partial[off + i]);
    off += nb[d];
  }
  for (int d = 0; d < 3; d++) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m[d] = dmax(m[d], __shfl_xor(m[d], o, 64));
    if ((threadIdx.x & 63) == 0) red[d][threadIdx.x >> 6] = m[d];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) / 64;
    double mm[3];
    for (int d = 0; d < 3; d++) {
      mm[d] = red[d][0];
      for (int w = 1; w < nw; w++) mm[d] = dmax(mm[d], red[d][w]);
    }
    // max over cells of dd*max(speeds) == dd*max(speeds) for a uniform grid
    double amax = 0.;
    if (a.nx > 1) amax = dmax(amax, a.ddx * mm[0]);
    if (a.ny > 1) amax = dmax(amax, a.ddy * mm[1]);
    if (a.nz > 1) amax = dmax(amax, a.ddz * mm[2]);
    if (a.res) {
      // resistive systems: the speed of light limits the step
      const double g3 = make_metric(a.coordinates, a.S->t).g3;
      amax = 0.;
      if (a.nx > 1) amax = dmax(amax, a.ddx);
      if (a.ny > 1) amax = dmax(amax, a.ddy);
      if (a.nz > 1) amax = dmax(amax, a.ddz / sqrt(g3));
    } else if (a.divclean) {
      amax = 1.;   // evolve.f90:734-736
    } else if (amax == 0) {
      if (a.nx > 1) amax = dmax(amax, a.ddx);
      if (a.ny > 1) amax = dmax(amax, a.ddy);
      if (a.nz > 1) amax = dmax(amax, a.ddz);
    }
    DevScalars* S = a.S;
    double dt_local = a.cfl / amax;
    dt_local = dmin(dt_local, a.maxdt);
    dt_local = dmin(dt_local, S->tout - S->t);
    if (fabs(S->t + S->dt - S->tout) < 1.e-9) dt_local = S->tout - S->t;
    else if ((S->t + 2. * dt_local - S->tout) > 1.e-9) dt_local = (S->tout - S->t) / 2.;
    S->dt = dt_local;
  }
}

// lagged time derivatives of u^mu (dts, viscous.f90:56-97) and the face
// differences of the reconstructed Lorentz velocities (evolve.f90:371-374)
__device__ __forceinline__ Deriv cell_deriv(const GridArgs& a, const Metric& M, long long c, int x,
                                            int y, int z, const double* __restrict__ LX,
                                            const double* __restrict__ LY,
                                            const double* __restrict__ LZ,
                                            const double* __restrict__ VSn,
                                            const double* __restrict__ VSo) {
  Deriv D;
  double dx[4] = {0., 0., 0., 0.}, dy[4] = {0., 0., 0., 0.}, dz[4] = {0., 0., 0., 0.};
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  if (nx > 1) {
    long long nf = (long long)(nx + 1) * ny * nz, f0 = x + (long long)(nx + 1) * (y + (long long)ny * z);
#pragma unroll
    for (int q = 0; q < 4; q++) dx[q] = a.ddx * 0.5 * (LX[q * nf + f0 + 1] - LX[q * nf + f0]);
  }
  if (ny > 1) {
    long long nf = (long long)nx * (ny + 1) * nz, f0 = x + (long long)nx * (y + (long long)(ny + 1) * z);
#pragma unroll
    for (int q = 0; q < 4; q++) dy[q] = a.ddy * 0.5 * (LY[q * nf + f0 + nx] - LY[q * nf + f0]);
  }
  if (nz > 1) {
    long long nf = (long long)nx * ny * (nz + 1), f0 = x + (long long)nx * (y + (long long)ny * z);
#pragma unroll
    for (int q = 0; q < 4; q++) dz[q] = a.ddz * 0.5 * (LZ[q * nf + f0 + (long long)nx * ny] - LZ[q * nf + f0]);
  }
  D.dutdx = dx[0]; D.duxdx = dx[1]; D.duydx = dx[2]; D.duzdx = dx[3];
  D.dutdy = dy[0]; D.duxdy = dy[1]; D.duydy = dy[2]; D.duzdy = dy[3];
  D.dutdz = dz[0]; D.duxdz = dz[1]; D.duydz = dz[2]; D.duzdz = dz[3];
  const DevScalars* S = a.S;
  const double ti = S->timeinterval;
  if (ti == 0) {
    D.dutdt = D.duxdt = D.duydt = D.duzdt = 0.;
  } else {
    const long long N = a.N;
    double ox = VSo[c], oy = VSo[N + c], oz = VSo[2 * N + c];
    double nxv = VSn[c], nyv = VSn[N + c], nzv = VSn[2 * N + c];
    double gamma1 = 1. / sqrt(1. - M.g1 * (ox * ox) - M.g2 * (oy * oy) - S->g_cov_3_old * (oz * oz));
    double gamma2 = 1. / sqrt(1. - M.g1 * (nxv * nxv) - M.g2 * (nyv * nyv) - M.g3 * (nzv * nzv));
    D.dutdt = (gamma2 - gamma1) / ti;
    D.duxdt = (gamma2 * nxv - gamma1 * ox) / ti;
    D.duydt = (gamma2 * nyv - gamma1 * oy) / ti;
    D.duzdt = (gamma2 * nzv - gamma1 * oz) / ti;
  }
  return D;
}

// div E of a resistive grid at a cell: each sweep stored, per face, the
// difference of E^d between the cells on either side (slot nv of its flux
// buffer); the central difference of system_compute_divE
// (system.f90:1901-1918) is the mean of the cell's two face differences.
__device__ __forceinline__ double res_div_e(const GridArgs& a, int x, int y, int z,
                                            const double* __restrict__ FX, const double* __restrict__ FY,
                                            const double* __restrict__ FZ, int nv) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  double r = 0.;
  if (nx > 1) {
    const long long nf = (long long)(nx + 1) * ny * nz, f0 = x + (long long)(nx + 1) * (y + (long long)ny * z);
    r = r + (FX[nv * nf + f0] + FX[nv * nf + f0 + 1]) * (a.ddx / 2.);
  }
  if (ny > 1) {
    const long long nf = (long long)nx * (ny + 1) * nz, f0 = x + (long long)nx * (y + (long long)(ny + 1) * z);
    r = r + (FY[nv * nf + f0] + FY[nv * nf + f0 + nx]) * (a.ddy / 2.);
  }
  if (nz > 1) {
    const long long nf = (long long)nx * ny * (nz + 1), f0 = x + (long long)nx * (y + (long long)ny * z);
    r = r + (FZ[nv * nf + f0] + FZ[nv * nf + f0 + (long long)nx * ny]) * (a.ddz / 2.);
  }
  return r;
}

// flux divergence of a cell, sum over the active directions of
// dd * (F(face+1) - F(face)) (evolve.f90:383 and the y/z analogues)
template <int NV>
__device__ __forceinline__ void flux_div(const GridArgs& a, int x, int y, int z, const double* __restrict__ FX,
                                         const double* __restrict__ FY, const double* __restrict__ FZ,
                                         double* du) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
#pragma unroll
  for (int k = 0; k < NV; k++) du[k] = 0.;
  if (nx > 1) {
    long long nf = (long long)(nx + 1) * ny * nz, f0 = x + (long long)(nx + 1) * (y + (long long)ny * z);
#pragma unroll
    for (int k = 0; k < NV; k++) du[k] = du[k] + a.ddx * (FX[k * nf + f0 + 1] - FX[k * nf + f0]);
  }
  if (ny > 1) {
    long long nf = (long long)nx * (ny + 1) * nz, f0 = x + (long long)nx * (y + (long long)(ny + 1) * z);
#pragma unroll
    for (int k = 0; k < NV; k++) du[k] = du[k] + a.ddy * (FY[k * nf + f0 + nx] - FY[k * nf + f0]);
  }
  if (nz > 1) {
    long long nf = (long long)nx * ny * (nz + 1), f0 = x + (long long)nx * (y + (long long)ny * z);
    const long long st = (long long)nx * ny;
#pragma unroll
    for (int k = 0; k < NV; k++) du[k] = du[k] + a.ddz * (FZ[k * nf + f0 + st] - FZ[k * nf + f0]);
  }
}

// ---------------------------------------------------------------------------
// k_update: one thread per cell.  Flux divergence (evolve.f90:383 and the
// y/z analogues) minus the sources (evolve_src, evolve.f90:689-718), the RK
// combination u = c1*u0 + c2*(u - dt*du) (evolve.f90:199), con2prim
// (evolve_u2v, evolve.f90:769-868) and, for viscous runs, the low-pressure
// smoothing (evolve_smooth_viscosity, evolve.f90:1159-1207).  Cells whose
// con2prim fails are marked in fb.failbits for k_fixup; what the repair
// reads of the other cells is kept in fb: every cell's pressure before the
// stage, and u before the smoothing of the smoothed cells.
// ---------------------------------------------------------------------------
template <int SYS, bool ZZ>
__global__ __launch_bounds__(UPDATE_BLOCK) void k_update(
    GridArgs a, double* __restrict__ V, double* U0, const double* Ucur, double* Uout,
    const double* __restrict__ FX, const double* __restrict__ FY, const double* __restrict__ FZ,
    const double* __restrict__ LX, const double* __restrict__ LY, const double* __restrict__ LZ,
    double* __restrict__ VS0, double* __restrict__ VS1, FixBufs fb, double c1,
    double c2, int first_stage, int last_stage) {
  constexpr bool VISC = SysT<SYS>::VISC;
  constexpr int NV = SysT<SYS>::NV;
  int x, y, z;
  long long c;
  if (!update_cell(a, x, y, z, c) || a.S->temp_check) return;
  DevScalars* S = a.S;
  if (S->run_crashed) return;
  const Metric M = make_metric(a.coordinates, S->t);
  const long long N = a.N;
  double du[NV];
  flux_div<NV>(a, x, y, z, FX, FY, FZ, du);
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = __builtin_nontemporal_load(&V[k * N + c]);   // read once
  __builtin_nontemporal_store(v[KPR], &fb.Pold[c]);
  Deriv D{};
  const int par = S->vs_parity;
  double* VSn = par ? VS1 : VS0;
  double* VSo = par ? VS0 : VS1;
  if constexpr (VISC) D = cell_deriv(a, M, c, x, y, z, LX, LY, LZ, VSn, VSo);
  {
    double vv[NV], src[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) vv[k] = v[k];
    int e;
    if constexpr (SysT<SYS>::RES) {
      // div E from the face differences the sweeps stored, and Ohm's
      // current taken explicitly (an explicit RK has no implicit part)
      const double rhoc = res_div_e(a, x, y, z, FX, FY, FZ, NV);
      double sst[3];
      e = res_source<SYS, ZZ>(a.eos, M, a.coordinates, a.vp, vv, D, rhoc, a.sig, a.sch, src, sst);
#pragma unroll
      for (int k = 0; k < 3; k++) src[SysT<SYS>::KE + k] = src[SysT<SYS>::KE + k] + sst[k];
    } else {
      e = sys_source<SYS, ZZ>(a.eos, M, a.coordinates, a.vp, vv, D, src);
    }
    if (e > 0) {
      atomicMax(&S->errcode, e);
      atomicExch(&S->run_crashed, 1);
      atomicCAS(&S->crash_kind, CRASH_NONE, CRASH_SOURCE);
      return;
    }
#pragma unroll
    for (int k = 0; k < NV; k++) du[k] = du[k] - src[k];
  }
  const double dt = S->dt;
  double u0[NV], uc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) {
    u0[k] = __builtin_nontemporal_load(&U0[k * N + c]);
    uc[k] = __builtin_nontemporal_load(&Ucur[k * N + c]);
  }
  if constexpr (SysT<SYS>::GLM) {
    constexpr int KG = SysT<SYS>::KG;
    if (first_stage) {
      // evolve_glm_source(dt/2) before u0 = u (evolve.f90:196): the damped
      // psi is also the RK base of the later stages
      const double ch = S->glm_ch;
      double ef = exp(-(ch * ch) * a.glm_alpha * (dt / 2.));
      u0[KG] = u0[KG] * ef;
      uc[KG] = u0[KG];
      U0[KG * N + c] = u0[KG];
    }
  }
  double un[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) un[k] = c1 * u0[k] + c2 * (uc[k] - dt * du[k]);
  if constexpr (SysT<SYS>::GLM) {
    constexpr int KG = SysT<SYS>::KG;
    if (last_stage) {   // evolve.f90:201
      const double ch = S->glm_ch;
      un[KG] = un[KG] * exp(-(ch * ch) * a.glm_alpha * (dt / 2.));
    }
  }
  if constexpr (SYS == SYS_MHD || SYS == SYS_MHDGLM) {
    if (a.ext_c2p) {
      // the learned inversion inverts u (k_mlp, then k_c2p_fallback_step
      // from the primitives V still holds)
#pragma unroll
      for (int k = 0; k < NV; k++) __builtin_nontemporal_store(un[k], &Uout[k * N + c]);
      return;
    }
  }
  int e = cons2prim<SYS, ZZ>(a.eos, M, un, v, a.algo_solver);
  if (e != 0) {
    atomicOr(&fb.failbits[c >> 5], 1u << (c & 31));
    atomicAdd(&S->nfail, 1);
  } else if (VISC && a.enable_smooth) {
    const double ps = S->smooth_pressure;
    if (v[KPR] < ps) {
#pragma unroll
      for (int k = 0; k < NV; k++) fb.Upre[k * N + c] = un[k];
      double q = v[KPR] / ps;
#pragma unroll
      for (int k = KPIBU; k <= KPIZZ; k++) v[k] = v[k] * (q * q);
      if constexpr (ZZ) derived_pi<true>(v, M);
      prim2cons<SYS, ZZ>(a.eos, M, v, un);
    }
  }
  // streamed outputs (read again only by the next kernel): non-temporal
  // stores, as in k_flux
#pragma unroll
  for (int k = 0; k < NV; k++) {
    __builtin_nontemporal_store(un[k], &Uout[k * N + c]);
    __builtin_nontemporal_store(v[k], &V[k * N + c]);
  }
  if (VISC && last_stage) {
    // velocities at the start of the next step (vnewstep, echo.f90:192)
    double* VSnext = VSo;
    __builtin_nontemporal_store(v[KVX], &VSnext[c]);
    __builtin_nontemporal_store(v[KVY], &VSnext[N + c]);
    __builtin_nontemporal_store(v[KVZ], &VSnext[2 * N + c]);
  }
}

// ---------------------------------------------------------------------------
// IMEX-SSP3(4,3,3) (Pareschi & Russo; coefficients of evolve.f90:127-130,
// stages of evolve.f90:204-264).  With du the
// explicit rate (flux divergence minus the non-stiff sources) and ds the
// stiff one (minus Ohm's current, only E), u = u0 - dt * (...):
//   K1 = u0                                  + a dt R(K1)
//   K2 = u0 + a dt ds1                       + a dt R(K2)
//   K3 = u0 - dt (du2 + (1 - a) ds2)         + a dt R(K3)
//   K4 = u0 - dt (du2/4 + du3/4 + b ds1 + h ds2 + d ds3) + a dt R(K4)
//   u1 = u0 - dt (du2/6 + du3/6 + 2/3 du4 + ds2/6 + ds3/6 + 2/3 ds4)
// (the explicit rate of K1 has weight zero, so the first stage needs no
// sweep).  Each implicit solve (res_cons2prim_implicit) inverts the stage
// start U* for the primitives and E; U keeps U*, V gets the stage's
// primitives.  The kernels below are shifted against the stages: each
// finishes the rates of one stage and solves the next stage's implicit part.
// On the other systems (hydro, viscous hydro, ideal MHD: no stiff term,
// ds = 0, as the reference runs INT_ALGO SSP there) every "implicit solve" is
// the plain inversion of the stage start, and the first stage sweeps only for
// maxval(aflux), the cleaning speed of the second (evolve.f90:225-231).
// ---------------------------------------------------------------------------
constexpr double IMEX_A = 0.24169426078821, IMEX_B = 0.06042356519705, IMEX_H = 0.12915286960590;
constexpr double IMEX_D = 0.5 - IMEX_B - IMEX_H - IMEX_A;

// implicit solve of one cell from its stage start us; marks a failure for
// k_fixup (which repairs with U* and, for the cells before it, V).  U* of
// the E components goes to Ust when STORE_E (the other components, and E in
// the first stage, the repair recomputes: ImexView).
template <int SYS, bool ZZ, bool STORE_E>
__device__ __forceinline__ void imex_solve(const GridArgs& a, const Metric& M, long long c, double* __restrict__ V,
                                           double* __restrict__ Ust, const FixBufs& fb, const double* us,
                                           double* v, double adt) {
  constexpr int NV = SysT<SYS>::NV;
  constexpr int KE = SysT<SYS>::KE;
  const long long N = a.N;
  __builtin_nontemporal_store(v[KPR], &fb.Pold[c]);
  double uu[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) {
    uu[k] = us[k];
    if (STORE_E && k >= KE && k < KE + 3) __builtin_nontemporal_store(us[k], &Ust[(long long)k * N + c]);
  }
  const int e = res_cons2prim_implicit<SYS, ZZ>(a.eos, M, uu, v, adt, a.sig, a.sch);
  if (e != 0) {
    atomicOr(&fb.failbits[c >> 5], 1u << (c & 31));
    atomicAdd(&a.S->nfail, 1);
  }
#pragma unroll
  for (int k = 0; k < NV; k++) __builtin_nontemporal_store(v[k], &V[(long long)k * N + c]);
}

// stiff rate ds = -sqrt(-g) R of a cell's primitives (E components), the
// du_stiff of evolve_src (evolve.f90:689-718)
template <int SYS>
__device__ __forceinline__ void imex_stiff_rate(const GridArgs& a, const Metric& M, const double* v, double* ds) {
  using T = SysT<SYS>;
  const double vel[3] = {v[KVX], v[KVY], v[KVZ]};
  double R[3];
  ohm_stiff(M, vel, &v[T::KE], &v[T::KB], a.sig, a.sch, R);
#pragma unroll
  for (int k = 0; k < 3; k++) ds[k] = -(R[k] * M.gp);
}

// stage 1: K1 from U* = u0
template <int SYS, bool ZZ>
__global__ __launch_bounds__(UPDATE_BLOCK) void k_imex_k1(GridArgs a, double* __restrict__ V,
                                                          const double* __restrict__ U0,
                                                          double* __restrict__ Ust, FixBufs fb) {
  constexpr int NV = SysT<SYS>::NV;
  const long long c = xcd_block(blockIdx.x, gridDim.x) * UPDATE_BLOCK + threadIdx.x;
  if (c >= a.N || a.S->temp_check || a.S->run_crashed) return;
  const long long N = a.N;
  const Metric M = make_metric(a.coordinates, a.S->t);
  double us[NV], v[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) {
    us[k] = U0[(long long)k * N + c];
    v[k] = V[(long long)k * N + c];
  }
  imex_solve<SYS, ZZ, false>(a, M, c, V, Ust, fb, us, v, IMEX_A * a.S->dt);
}

// the stiff rate of K1 (cell-local), P4 = b ds1, and K2 from
// U* = u0 + a dt ds1
template <int SYS, bool ZZ>
__global__ __launch_bounds__(UPDATE_BLOCK) void k_imex_k2(GridArgs a, double* __restrict__ V,
                                                          const double* __restrict__ U0,
                                                          double* __restrict__ Ust, ImexBufs ib, FixBufs fb) {
  using T = SysT<SYS>;
  constexpr int NV = T::NV;
  const long long c = xcd_block(blockIdx.x, gridDim.x) * UPDATE_BLOCK + threadIdx.x;
  if (c >= a.N || a.S->temp_check || a.S->run_crashed) return;
  const long long N = a.N;
  const Metric M = make_metric(a.coordinates, a.S->t);
  const double dt = a.S->dt;
  double us[NV], v[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) {
    us[k] = U0[(long long)k * N + c];
    v[k] = V[(long long)k * N + c];
  }
  double ds[3];
  imex_stiff_rate<SYS>(a, M, v, ds);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    ib.P4[(long long)(T::KE + k) * N + c] = IMEX_B * ds[k];
    us[T::KE + k] = us[T::KE + k] + dt * IMEX_A * ds[k];
  }
  imex_solve<SYS, ZZ, true>(a, M, c, V, Ust, fb, us, v, IMEX_A * dt);
}

// systems without a stiff term, second stage: V from u0 (evolve.f90:232-233,
// u = u0 + a dt ds1 with ds1 = 0); U keeps u0
template <int SYS, bool ZZ>
__global__ __launch_bounds__(UPDATE_BLOCK) void k_imex_u2v(GridArgs a, double* __restrict__ V,
                                                           const double* __restrict__ U0, FixBufs fb) {
  constexpr int NV = SysT<SYS>::NV;
  const long long c = xcd_block(blockIdx.x, gridDim.x) * UPDATE_BLOCK + threadIdx.x;
  if (c >= a.N || a.S->temp_check || a.S->run_crashed) return;
  const long long N = a.N;
  const Metric M = make_metric(a.coordinates, a.S->t);
  double u[NV], v[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) {
    u[k] = U0[(long long)k * N + c];
    v[k] = V[(long long)k * N + c];
  }
  __builtin_nontemporal_store(v[KPR], &fb.Pold[c]);
  const int e = cons2prim<SYS, ZZ>(a.eos, M, u, v, a.algo_solver);
  if (e != 0) {
    atomicOr(&fb.failbits[c >> 5], 1u << (c & 31));
    atomicAdd(&a.S->nfail, 1);
  }
#pragma unroll
  for (int k = 0; k < NV; k++) __builtin_nontemporal_store(v[k], &V[(long long)k * N + c]);
}

// stages 2-4 after their sweeps: the rates of stage STAGE at K_STAGE
// (flux divergence, non-stiff sources with the charge density div E, stiff
// rate), their accumulation, and K_{STAGE+1} (STAGE 2, 3) or the step's
// result with the GLM damping of psi (evolve.f90:262), its plain inversion
// and the viscous smoothing (STAGE 4)
template <int SYS, bool ZZ, int STAGE>
__global__ __launch_bounds__(UPDATE_BLOCK) void k_update_imex(
    GridArgs a, double* __restrict__ V, double* __restrict__ U0, double* __restrict__ Ust,
    const double* __restrict__ FX, const double* __restrict__ FY, const double* __restrict__ FZ,
    const double* __restrict__ LX, const double* __restrict__ LY, const double* __restrict__ LZ,
    double* __restrict__ VS0, double* __restrict__ VS1, ImexBufs ib, FixBufs fb) {
  using T = SysT<SYS>;
  constexpr bool VISC = T::VISC;
  constexpr int NV = T::NV;
  int x, y, z;
  long long c;
  if (!update_cell(a, x, y, z, c) || a.S->temp_check) return;
  DevScalars* S = a.S;
  if (S->run_crashed) return;
  const Metric M = make_metric(a.coordinates, S->t);
  const long long N = a.N;
  double du[NV];
  flux_div<NV>(a, x, y, z, FX, FY, FZ, du);
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = __builtin_nontemporal_load(&V[k * N + c]);
  Deriv D{};
  const int par = S->vs_parity;
  double* VSn = par ? VS1 : VS0;
  double* VSo = par ? VS0 : VS1;
  if constexpr (VISC) D = cell_deriv(a, M, c, x, y, z, LX, LY, LZ, VSn, VSo);
  double ds[3] = {0., 0., 0.};
  {
    double vv[NV], src[NV], sst[3] = {0., 0., 0.};
#pragma unroll
    for (int k = 0; k < NV; k++) vv[k] = v[k];
    int e;
    if constexpr (T::RES) {
      const double rhoc = res_div_e(a, x, y, z, FX, FY, FZ, NV);
      e = res_source<SYS, ZZ>(a.eos, M, a.coordinates, a.vp, vv, D, rhoc, a.sig, a.sch, src, sst);
    } else {
      e = sys_source<SYS, ZZ>(a.eos, M, a.coordinates, a.vp, vv, D, src);
    }
    if (e > 0) {
      atomicMax(&S->errcode, e);
      atomicExch(&S->run_crashed, 1);
      atomicCAS(&S->crash_kind, CRASH_NONE, CRASH_SOURCE);
      return;
    }
#pragma unroll
    for (int k = 0; k < NV; k++) du[k] = du[k] - src[k];
#pragma unroll
    for (int k = 0; k < 3; k++) ds[k] = -sst[k];
  }
  const double dt = S->dt;
  double u0[NV], un[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) u0[k] = __builtin_nontemporal_load(&U0[k * N + c]);
  // u = fma(-dt, rate, u0); the repair recomputes U* of the components other
  // than E from P4 with the same operation (ImexView)
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const int ke = k - T::KE;
    const bool isE = T::RES && ke >= 0 && ke < 3;
    const long long o = (long long)k * N + c;
    if (isE) {
      const double s = ds[ke];
      if constexpr (STAGE == 2) {
        ib.P4[o] = ib.P4[o] + (du[k] / 4. + IMEX_H * s);
        ib.ACC[o] = du[k] / 6. + s / 6.;
        un[k] = __builtin_fma(-dt, du[k] + (1. - IMEX_A) * s, u0[k]);
      } else if constexpr (STAGE == 3) {
        un[k] = __builtin_fma(-dt, ib.P4[o] + du[k] / 4. + IMEX_D * s, u0[k]);
        ib.ACC[o] = ib.ACC[o] + (du[k] / 6. + s / 6.);
      } else {
        un[k] = __builtin_fma(-dt, ib.ACC[o] + du[k] * 2. / 3. + s * 2. / 3., u0[k]);
      }
    } else {
      if constexpr (STAGE == 2) {
        ib.P4[o] = du[k] * 0.25;
        un[k] = __builtin_fma(-dt, du[k], u0[k]);
      } else if constexpr (STAGE == 3) {
        const double r3 = ib.P4[o] + du[k] * 0.25;
        ib.P4[o] = r3;
        un[k] = __builtin_fma(-dt, r3, u0[k]);
      } else {
        un[k] = __builtin_fma(-dt, ib.P4[o] * (2. / 3.) + du[k] * 2. / 3., u0[k]);
      }
    }
  }
  if constexpr (STAGE < 4 && T::RES) {
    imex_solve<SYS, ZZ, true>(a, M, c, V, Ust, fb, un, v, IMEX_A * dt);
  } else if constexpr (STAGE < 4) {
    // no stiff term: the plain inversion of K_{STAGE+1} (evolve.f90:241,250)
    __builtin_nontemporal_store(v[KPR], &fb.Pold[c]);
    const int e = cons2prim<SYS, ZZ>(a.eos, M, un, v, a.algo_solver);
    if (e != 0) {
      atomicOr(&fb.failbits[c >> 5], 1u << (c & 31));
      atomicAdd(&S->nfail, 1);
    }
#pragma unroll
    for (int k = 0; k < NV; k++) __builtin_nontemporal_store(v[k], &V[(long long)k * N + c]);
  } else {
    if constexpr (T::GLM) {
      const double ch = S->glm_ch;
      un[T::KG] = un[T::KG] * exp(-(ch * ch) * a.glm_alpha * (dt / 2.));
    }
    __builtin_nontemporal_store(v[KPR], &fb.Pold[c]);
    const int e = cons2prim<SYS, ZZ>(a.eos, M, un, v, a.algo_solver);
    if (e != 0) {
      atomicOr(&fb.failbits[c >> 5], 1u << (c & 31));
      atomicAdd(&S->nfail, 1);
    } else if (VISC && a.enable_smooth) {
      const double ps = S->smooth_pressure;
      if (v[KPR] < ps) {
#pragma unroll
        for (int k = 0; k < NV; k++) fb.Upre[k * N + c] = un[k];
        double q = v[KPR] / ps;
#pragma unroll
        for (int k = KPIBU; k <= KPIZZ; k++) v[k] = v[k] * (q * q);
        if constexpr (ZZ) derived_pi<true>(v, M);
        prim2cons<SYS, ZZ>(a.eos, M, v, un);
      }
    }
#pragma unroll
    for (int k = 0; k < NV; k++) {
      __builtin_nontemporal_store(un[k], &U0[k * N + c]);
      __builtin_nontemporal_store(v[k], &V[k * N + c]);
    }
    if (VISC) {
      __builtin_nontemporal_store(v[KVX], &VSo[c]);
      __builtin_nontemporal_store(v[KVY], &VSo[N + c]);
      __builtin_nontemporal_store(v[KVZ], &VSo[2 * N + c]);
    }
  }
}

// ---------------------------------------------------------------------------
// k_fixup: the con2prim repair of evolve_u2v (evolve.f90:795-842) for the
// cells that k_update could not invert, with the reference's sequential
// semantics.  The reference repairs cell by cell in loop order (x fastest)
// while it inverts the grid, so a failed cell sees:
//   - u of its face neighbours after their inversion (and repair) when they
//     come earlier in the loop, the RK result when they come later; both
//     before the viscous smoothing, which runs after the whole loop;
//   - the pressure of its neighbours from the new primitives when they come
//     earlier, from the old ones (before the stage) when they come later.
// Here the failed cells are taken in dependency levels: a failed cell waits
// for its earlier failed face neighbours (lower levels), and is done before
// its later ones (higher levels), so within a level the cells are
// independent and run in parallel.  The neighbour values come from U and V
// (new), fb.Pold (old pressure) and fb.Upre (u of smoothed cells before the
// smoothing).  Boundary neighbours are clamped to the cell itself (the
// reference reads outside its arrays there).  The smoothing of the repaired
// cells follows the repairs.  One block.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool fix_failed(const unsigned* __restrict__ bits, long long q) {
  return (bits[q >> 5] >> (q & 31)) & 1u;
}

// U* of component k of cell q (ImexView) as the repair reads it; repaired:
// q was repaired before the cell being repaired (its whole u is in U)
template <int SYS>
__device__ __forceinline__ double imex_ustar(const ImexView& iv, const double* __restrict__ U, bool repaired,
                                             long long q, int k, long long N, double dt) {
  constexpr int KE = SysT<SYS>::KE;
  const long long o = (long long)k * N + q;
  const bool isE = SysT<SYS>::RES && k >= KE && k < KE + 3;
  if (repaired || (isE && iv.phase > 1)) return U[o];
  if (iv.phase <= 2) return iv.U0[o];
  const double r = iv.phase == 3 ? 4.0 * iv.P4[o] : iv.P4[o];
  return __builtin_fma(-dt, r, iv.U0[o]);
}

template <int SYS, bool ZZ>
__device__ void repair_cell(const GridArgs& a, const Metric& M, long long c, double* __restrict__ V,
                            double* __restrict__ U, const FixBufs& fb, double* __restrict__ VSnext,
                            int last_stage, const ImexView& iv, int smooth_stage) {
  constexpr bool VISC = SysT<SYS>::VISC;
  constexpr int NV = SysT<SYS>::NV;
  DevScalars* S = a.S;
  const long long N = a.N;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int x = (int)(c % nx);
  const long long r = c / nx;
  const int y = (int)(r % ny), z = (int)(r / ny);
  const long long sxy = (long long)nx * ny;
  const long long nb[6] = {c - (x > 0 ? 1 : 0), c + (x < nx - 1 ? 1 : 0),
                           c - (y > 0 ? nx : 0), c + (y < ny - 1 ? nx : 0),
                           c - (z > 0 ? sxy : 0), c + (z < nz - 1 ? sxy : 0)};
  // u of a neighbour as the reference's loop sees it: a smoothed cell's u
  // before its smoothing (fb.Upre), else U (the RK result of later cells,
  // failed ones included, and the repaired u of earlier failed cells).
  // Inside an implicit IMEX stage a neighbour's u is its explicit start U*
  // (imex_ustar; the repaired u for earlier failed cells), and V holds the
  // E its implicit solve found: earlier cells are read with that E (u_E =
  // sqrt(-g) E), later ones (not yet solved in the reference's loop) with E*.
  const int imex = iv.phase;
  const double dt = S->dt;
  bool repaired[6];
#pragma unroll
  for (int j = 0; j < 6; j++) repaired[j] = nb[j] < c && fix_failed(fb.failbits, nb[j]);
  const double* us[6];
  const bool smooth = VISC && a.enable_smooth && smooth_stage;
  const double ps = S->smooth_pressure;
#pragma unroll
  for (int j = 0; j < 6; j++) {
    const long long q = nb[j];
    const bool pre = smooth && q != c && !fix_failed(fb.failbits, q) && V[KPR * N + q] < ps;
    us[j] = pre ? fb.Upre : U;
  }
  double uu[NV], vv[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const long long o = (long long)k * N;
    double w[6];
#pragma unroll
    for (int j = 0; j < 6; j++) w[j] = us[j][o + nb[j]];
    if (imex) {
#pragma unroll
      for (int j = 0; j < 6; j++) w[j] = imex_ustar<SYS>(iv, U, repaired[j], nb[j], k, N, dt);
    }
    if constexpr (SysT<SYS>::RES) {
      constexpr int KE = SysT<SYS>::KE;
      if (imex && k >= KE && k < KE + 3) {
#pragma unroll
        for (int j = 0; j < 6; j++)
          if (nb[j] < c) w[j] = V[o + nb[j]] * M.gp;
      }
    }
    uu[k] = (w[0] + w[1] + w[2] + w[3] + w[4] + w[5]) / 6.;
    vv[k] = V[o + c];   // the primitives k_update's failed inversion left
  }
  cons2prim<SYS, ZZ>(a.eos, M, uu, vv, a.algo_solver);
  double v2 = vv[KVX] * vv[KVX] + vv[KVY] * vv[KVY] + vv[KVZ] * vv[KVZ] * M.g3;
  if (v2 > 1) {
    const double sc = a.maxspeed / sqrt(v2);
    vv[KVX] = sc * vv[KVX]; vv[KVY] = sc * vv[KVY]; vv[KVZ] = sc * vv[KVZ];
  }
  if (vv[KPR] < 0) {
    double p[6];
#pragma unroll
    for (int j = 0; j < 6; j++) p[j] = nb[j] < c ? V[KPR * N + nb[j]] : fb.Pold[nb[j]];
    vv[KPR] = (p[0] + p[1] + p[2] + p[3] + p[4] + p[5]) / 6.;
  }
  prim2cons<SYS, ZZ>(a.eos, M, vv, uu);
  cons2prim<SYS, ZZ>(a.eos, M, uu, vv, a.algo_solver);
  v2 = vv[KVX] * vv[KVX] + vv[KVY] * vv[KVY] + vv[KVZ] * vv[KVZ] * M.g3;
  // u is stored before the check (evolve.f90:848), pi^{eta eta} too with 'zz'
#pragma unroll
  for (int k = 0; k < NV; k++) U[(long long)k * N + c] = uu[k];
  if constexpr (VISC && ZZ) V[KPIZZ * N + c] = vv[KPIZZ];
  if (!((vv[KPR] > 0) && (v2 < 1))) {
    atomicMax(&S->errcode, 1);
    atomicExch(&S->run_crashed, 1);
    atomicCAS(&S->crash_kind, CRASH_NONE, CRASH_REPAIR);
    atomicMin(&S->crash_cell, (int)c);
    return;
  }
#pragma unroll
  for (int k = 0; k < NV; k++) V[(long long)k * N + c] = vv[k];
  if (VISC && last_stage) {
    VSnext[c] = vv[KVX]; VSnext[N + c] = vv[KVY]; VSnext[2 * N + c] = vv[KVZ];
  }
}

template <int SYS, bool ZZ>
__global__ __launch_bounds__(FIX_BLOCK) void k_fixup(GridArgs a, double* __restrict__ V, double* __restrict__ U,
                                                     FixBufs fb, double* __restrict__ VS0,
                                                     double* __restrict__ VS1, int last_stage, ImexView iv,
                                                     int smooth_stage) {
  constexpr bool VISC = SysT<SYS>::VISC;
  constexpr int NV = SysT<SYS>::NV;
  DevScalars* S = a.S;
  const int t = threadIdx.x;
  __shared__ int sh_go, sh_maxlev, sh_stop;
  __shared__ int scan[FIX_BLOCK + 1];
  if (t == 0) {
    sh_go = 0;
    if (!S->temp_check) {
      // glm_ch_old = maxval(aflux) of this stage; the next stage starts with
      // glm_ch = glm_ch_old (evolve.f90:157-163)
      // (resistive systems clean at the speed of light)
      S->glm_ch_old = a.res ? 1. : dmax(dmax(S->aflux[0], S->aflux[1]), S->aflux[2]);
      S->glm_ch = S->glm_ch_old;
      sh_go = S->nfail > 0;
    }
  }
  __syncthreads();
  if (!sh_go) return;
  const long long N = a.N;
  const long long words = (N + 31) >> 5;
  const long long per = (words + FIX_BLOCK - 1) / FIX_BLOCK;
  const long long w0 = t * per < words ? t * per : words;
  const long long w1 = w0 + per < words ? w0 + per : words;
  // 1. the failed cells in increasing order: popcounts per thread chunk,
  //    exclusive scan, then each thread lists its own chunk
  int cnt = 0;
  for (long long w = w0; w < w1; w++) cnt += __popc(fb.failbits[w]);
  scan[t + 1] = cnt;
  __syncthreads();
  if (t == 0) {
    scan[0] = 0;
    for (int i = 1; i <= FIX_BLOCK; i++) scan[i] += scan[i - 1];
  }
  __syncthreads();
  const int n = scan[FIX_BLOCK];
  {
    int o = scan[t];
    for (long long w = w0; w < w1; w++) {
      unsigned b = fb.failbits[w];
      while (b) {
        const int i = __ffs(b) - 1;
        b &= b - 1;
        fb.list[o++] = (int)(w * 32 + i);
      }
    }
  }
  __syncthreads();
  // 2. dependency levels: one more than the highest level of the failed face
  //    neighbours that come earlier (x-1, y-1, z-1), i.e. the longest chain
  //    of earlier failed neighbours ending at the cell.  All threads relax
  //    the levels from zero until a pass changes none; the values only grow
  //    and never pass the fixed point, which the pass that changes nothing
  //    has reached (at most one pass per level, plus one).
  const int nx = a.nx, ny = a.ny;
  const long long sxy = (long long)nx * ny;
  for (int i = t; i < n; i += FIX_BLOCK) fb.level[fb.list[i]] = 0;
  if (t == 0) { sh_maxlev = 0; sh_stop = 0; }
  __syncthreads();
  int mylev = 0;
  for (;;) {
    int changed = 0;
    for (int i = t; i < n; i += FIX_BLOCK) {
      const long long c = fb.list[i];
      const int x = (int)(c % nx);
      const long long r = c / nx;
      const int y = (int)(r % ny), z = (int)(r / ny);
      int lev = 0;
      if (x > 0 && fix_failed(fb.failbits, c - 1)) lev = max(lev, fb.level[c - 1] + 1);
      if (y > 0 && fix_failed(fb.failbits, c - nx)) lev = max(lev, fb.level[c - nx] + 1);
      if (z > 0 && fix_failed(fb.failbits, c - sxy)) lev = max(lev, fb.level[c - sxy] + 1);
      if (lev > fb.level[c]) {
        fb.level[c] = lev;
        changed = 1;
      }
      mylev = max(mylev, lev);
    }
    if (!__syncthreads_or(changed)) break;
  }
  atomicMax(&sh_maxlev, mylev);
  __syncthreads();
  // 3. the repairs, one level after the other.  A failed repair crashes the
  //    run; the reference returns from evolve_u2v there, leaving the later
  //    cells as they were, while here every level is still repaired: the
  //    cells before the crashing one end as in the reference, the state of
  //    the later ones is not meaningful after a crash either way.
  const Metric M = make_metric(a.coordinates, S->t);
  double* VSnext = S->vs_parity ? VS0 : VS1;
  for (int lev = 0; lev <= sh_maxlev; lev++) {
    for (int i = t; i < n; i += FIX_BLOCK) {
      const long long c = fb.list[i];
      if (fb.level[c] == lev) repair_cell<SYS, ZZ>(a, M, c, V, U, fb, VSnext, last_stage, iv, smooth_stage);
    }
    __syncthreads();
  }
  if (t == 0) sh_stop = S->run_crashed;
  __syncthreads();
  // 4. evolve_smooth_viscosity (evolve.f90:1159-1207) of the repaired cells
  //    (not after a crash: evolve_main returns before it, evolve.f90:276-281)
  if (VISC && a.enable_smooth && smooth_stage && !sh_stop) {
    const double ps = S->smooth_pressure;
    for (int i = t; i < n; i += FIX_BLOCK) {
      const long long c = fb.list[i];
      if (V[KPR * N + c] < ps) {
        double vv[NV], uu[NV];
#pragma unroll
        for (int k = 0; k < NV; k++) vv[k] = V[(long long)k * N + c];
        const double q = vv[KPR] / ps;
#pragma unroll
        for (int k = KPIBU; k <= KPIZZ; k++) vv[k] = vv[k] * (q * q);
        if constexpr (ZZ) derived_pi<true>(vv, M);
        prim2cons<SYS, ZZ>(a.eos, M, vv, uu);
#pragma unroll
        for (int k = 0; k < NV; k++) { U[(long long)k * N + c] = uu[k]; V[(long long)k * N + c] = vv[k]; }
      }
    }
  }
  // 5. clear the marks for the next stage
  for (long long w = w0; w < w1; w++)
    if (fb.failbits[w]) fb.failbits[w] = 0u;
  if (t == 0) {
    S->nfail = 0;
    S->nrepair += n;
  }
}

// end of step bookkeeping, echo.f90:211-247 (time, metric, dts interval,
// vnewstep -> vold, tout cadence)
// vnewstep / unewstep of echo.f90:192-201 for the restart dump: components
// krh..kvold_end of v and u at the start of the step go to the buffer of the
// step's parity (iter & 1); between steps the buffer of parity
// (iter - 1) & 1 is vold / uold (echo.f90:240-241).  Each buffer holds the
// kend v components, then the kend u components.
static __global__ __launch_bounds__(UPDATE_BLOCK) void k_dump_copy(const DevScalars* S,
                                                                  const double* __restrict__ V,
                                                                  const double* __restrict__ U,
                                                                  double* __restrict__ B0,
                                                                  double* __restrict__ B1, long long n) {
  if (S->temp_check || S->run_crashed) return;
  const long long i = (long long)blockIdx.x * UPDATE_BLOCK + threadIdx.x;
  if (i >= n) return;
  double* B = (S->iter & 1) ? B1 : B0;
  B[i] = V[i];
  B[n + i] = U[i];
}

static __global__ void k_end_step(DevScalars* S, int coordinates, int flip_vs) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (S->run_crashed || S->temp_check) return;
  S->timeold = S->t;
  S->t = S->t + S->dt;
  // g_cov_3_old = g_cov(3) of the step just taken (echo.f90:213)
  S->g_cov_3_old = make_metric(coordinates, S->timeold).g3;
  S->iter = S->iter + 1;
  S->timeinterval = S->dt;
  if (flip_vs) S->vs_parity = S->vs_parity ^ 1;
  if ((S->t > S->tout) || ((S->tout - S->t) < 1.e-10)) {
    S->tout = S->tout + S->dtout;
    if ((fabs(S->tout - S->tmax) < 1.e-10) || (S->tout > S->tmax)) S->tout = S->tmax;
  }
}

// Output frames (out_vars, out.f90:232-591): the derived fields of a
// frame, per cell.  dst[j] is the frame slot of derived field j or -1:
// 0 energy density, 1 temperature, 2 entropy density (get_derived_data),
// 3..6 pi^{tt}, pi^{tx}, pi^{ty}, pi^{t eta} and 7 pi^{eta eta} (TRACE_IMP
// 'zz') from get_derived_pi(_zz), 8 gamma, 9..11 the ideal-MHD electric
// field of system_compute_E_in_ideal_MHD (system.f90:1885-1897) with E_eta
// made contravariant.  The frame is [t][arrays], slot s at 1 + s*N.
struct OutSlots { int dst[12]; };

template <int SYS, bool ZZ>
__global__ __launch_bounds__(UPDATE_BLOCK) void k_out_derived(GridArgs a, const double* __restrict__ V,
                                                             OutSlots o, double* __restrict__ frame,
                                                             int* __restrict__ status) {
  constexpr int NV = SysT<SYS>::NV;
  const long long c = (long long)blockIdx.x * UPDATE_BLOCK + threadIdx.x;
  if (c >= a.N) return;
  const long long N = a.N;
  const Metric M = make_metric(a.coordinates, a.S->t);
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = V[k * N + c];
  double* F = frame + 1;
  if (o.dst[0] >= 0 || o.dst[1] >= 0 || o.dst[2] >= 0) {
    double en = 0., temp = 0., s = 0.;
    const int err = get_derived_data(a.eos, v[KRH], v[KPR], en, temp, s);
    if (err) atomicMax(status, err);
    if (o.dst[0] >= 0) F[o.dst[0] * N + c] = en;
    if (o.dst[1] >= 0) F[o.dst[1] * N + c] = temp;
    if (o.dst[2] >= 0) F[o.dst[2] * N + c] = s;
  }
  if constexpr (SysT<SYS>::VISC) {
    if (o.dst[3] >= 0 || o.dst[4] >= 0 || o.dst[5] >= 0 || o.dst[6] >= 0 || o.dst[7] >= 0) {
      const PiDer d = derived_pi<ZZ>(v, M);
      if (o.dst[3] >= 0) F[o.dst[3] * N + c] = d.pitt;
      if (o.dst[4] >= 0) F[o.dst[4] * N + c] = d.pitx;
      if (o.dst[5] >= 0) F[o.dst[5] * N + c] = d.pity;
      if (o.dst[6] >= 0) F[o.dst[6] * N + c] = d.pitz;
      if (o.dst[7] >= 0) F[o.dst[7] * N + c] = v[KPIZZ];
    }
  }
  if (o.dst[8] >= 0) {
    const double vx = v[KVX], vy = v[KVY], vz = v[KVZ];
    F[o.dst[8] * N + c] = 1. / sqrt(1. - (vx * vx) * M.g1 - (vy * vy) * M.g2 - (vz * vz) * M.g3);
  }
  if constexpr (SysT<SYS>::MHD) {
    const double vx = v[KVX], vy = v[KVY], vz = v[KVZ], bx = v[KBX], by = v[KBY], bz = v[KBZ];
    if (o.dst[9] >= 0) F[o.dst[9] * N + c] = -M.gp * (vy * bz - vz * by);
    if (o.dst[10] >= 0) F[o.dst[10] * N + c] = -M.gp * (vz * bx - vx * bz);
    if (o.dst[11] >= 0) F[o.dst[11] * N + c] = -M.gp * (vx * by - vy * bx) / M.g3;
  }
}

// check_temp (out.f90:1487-1557), run at the start of each step
// (echo.f90:189-190): the maximum over the cells of the temperature
// (freeze_type 0) or energy density (1) from get_derived_temp_for_check
// (eos.f90:706-759; the same interpolation as get_derived_data).
// k_check_temp writes one maximum per block, k_check_decide reduces them and
// sets temp_check when the maximum is below temp_end; every later kernel of
// the run then returns at once, as the reference leaves its time loop.
constexpr int CHECK_BLOCK = 256;

static __global__ __launch_bounds__(CHECK_BLOCK) void k_check_temp(GridArgs a, const double* __restrict__ V,
                                                                   double* __restrict__ partial) {
  double m = 0.;
  const long long c = (long long)blockIdx.x * CHECK_BLOCK + threadIdx.x;
  if (c < a.N && !a.S->temp_check) {
    double en, temp, s;
    get_derived_data(a.eos, V[c], V[KPR * a.N + c], en, temp, s);
    m = a.S->freeze_type == 0 ? temp : en;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = dmax(m, __shfl_xor(m, o, 64));
  __shared__ double red[CHECK_BLOCK / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < CHECK_BLOCK / 64; w++) r = dmax(r, red[w]);
    partial[blockIdx.x] = r;
  }
}

static __global__ __launch_bounds__(1024) void k_check_decide(DevScalars* S, const double* __restrict__ partial,
                                                              int nb) {
  if (S->temp_check) return;
  __shared__ double red[16];
  double m = 0.;
  for (int i = threadIdx.x; i < nb; i += 1024) m = dmax(m, partial[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = dmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < 16; w++) r = dmax(r, red[w]);
    S->check_max = r;
    if (r < S->temp_end) S->temp_check = 1;
  }
}

// initial state: per-cell prim2cons then cons2prim (init.f90:2937-2958)
template <int SYS, bool ZZ>
__global__ __launch_bounds__(UPDATE_BLOCK) void k_init(GridArgs a, double* __restrict__ V,
                                                       double* __restrict__ U, double* __restrict__ VS0,
                                                       int* __restrict__ status) {
  constexpr bool VISC = SysT<SYS>::VISC;
  constexpr int NV = SysT<SYS>::NV;
  const long long c = (long long)blockIdx.x * UPDATE_BLOCK + threadIdx.x;
  if (c >= a.N) return;
  const long long N = a.N;
  const Metric M = make_metric(a.coordinates, a.S->t);
  double v[NV], u[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = V[k * N + c];
  int e = prim2cons<SYS, ZZ>(a.eos, M, v, u);
  if (e == 0 && !VISC) e = cons2prim<SYS, ZZ>(a.eos, M, u, v, a.algo_solver);
  if (e) atomicMax(status, e);
#pragma unroll
  for (int k = 0; k < NV; k++) { V[k * N + c] = v[k]; U[k * N + c] = u[k]; }
  if (VISC) { VS0[c] = v[KVX]; VS0[N + c] = v[KVY]; VS0[2 * N + c] = v[KVZ]; }
}

// ---------------------------------------------------------------------------
// Glauber-MC initial state.  k_glauber_sums: per transverse cell, the sums
// of the Gaussian sources of the participants of nuclei A and B and of the
// binary collisions around the centre of mass (glaubermc.f90:822-842); the
// sources are staged in LDS.  k_glauber_prim: per cell, the energy density
// (those sums with the (ybeam +- eta)/ybeam weights, kappa, ah and the
// longitudinal profile feta, glaubermc.f90:661-682,838), the primitives of
// the Glauber-MC initialisation (init.f90:2662-2730) and, for viscous runs,
// the Navier-Stokes stresses of viscous_initio (viscous.f90:101-167).
// ---------------------------------------------------------------------------
struct GlauberArgs {
  double kappa, ah, s2, ybeam, deta, sigeta, rhocenter, cmx, cmy;
  int na, nb, nc;
};

constexpr int GLAUBER_BLOCK = 256;
constexpr int GLAUBER_MAX_SRC = 4096;   // LDS: 64 KiB of (x, y) pairs

static __global__ __launch_bounds__(GLAUBER_BLOCK) void k_glauber_sums(GridArgs a, GlauberArgs q, double x1,
                                                               double dx, double y1, double dy,
                                                               const double* __restrict__ src,
                                                               double* __restrict__ sums) {
  __shared__ double2 s[GLAUBER_MAX_SRC];
  const int ns = q.na + q.nb + q.nc;
  for (int i = threadIdx.x; i < ns; i += GLAUBER_BLOCK) s[i] = make_double2(src[2 * i], src[2 * i + 1]);
  __syncthreads();
  const int c = blockIdx.x * GLAUBER_BLOCK + threadIdx.x;
  if (c >= a.nx * a.ny) return;
  const double x = x1 + ((double)(c % a.nx + 1) - 0.5) * dx;
  const double y = y1 + ((double)(c / a.nx + 1) - 0.5) * dy;
  double e[3] = {0., 0., 0.};
  int i = 0;
#pragma unroll
  for (int part = 0; part < 3; part++) {
    const int end = i + (part == 0 ? q.na : part == 1 ? q.nb : q.nc);
    double acc = 0.;
    for (; i < end; i++) {
      const double u = x - s[i].x + q.cmx, w = y - s[i].y + q.cmy;
      acc = acc + exp(-(u * u + w * w) / 2. / q.s2);
    }
    e[part] = acc;
  }
  const long long n2 = (long long)a.nx * a.ny;
  sums[c] = e[0];
  sums[n2 + c] = e[1];
  sums[2 * n2 + c] = e[2];
}

template <int SYS>
__global__ __launch_bounds__(UPDATE_BLOCK) void k_glauber_prim(GridArgs a, GlauberArgs q, double z1, double dz,
                                                               double tau0, double enezero, double pi,
                                                               const double* __restrict__ sums,
                                                               double* __restrict__ V, int* __restrict__ status) {
  constexpr int NV = SysT<SYS>::NV;
  const long long c = (long long)blockIdx.x * UPDATE_BLOCK + threadIdx.x;
  if (c >= a.N) return;
  const long long n2 = (long long)a.nx * a.ny;
  const long long t = c % n2;
  const int iz = (int)(c / n2);
  const double z = z1 + ((double)(iz + 1) - 0.5) * dz;
  const double ea = sums[t], eb = sums[n2 + t], ec = sums[2 * n2 + t];
  const double ep = ea * (q.ybeam + z) / q.ybeam + eb * (q.ybeam - z) / q.ybeam;
  double fe;
  const double half = q.deta / 2.;
  if (fabs(z) >= q.ybeam) fe = 0.;
  else if (fabs(z) <= half) fe = 1.;
  else {
    const double d = fabs(z) - half;
    fe = exp(-(d * d) / (2. * (q.sigeta * q.sigeta)));
  }
  const double e = q.kappa / (2. * pi * q.s2) * (q.ah * ec + (1 - q.ah) * ep) * fe;
  double pr, d1, d2;
  int err = eos_pressure(a.eos, q.rhocenter, e + enezero, pr, d1, d2);
#pragma unroll
  for (int k = 0; k < NV; k++) V[k * a.N + c] = 0.;
  V[KRH * a.N + c] = 1.;
  V[KPR * a.N + c] = pr;
  if constexpr (SysT<SYS>::VISC) {
    double en, temp, s, cs2;
    if (err == 0) err = get_derived_data(a.eos, 1., pr, en, temp, s);
    if (err == 0) err = eos_sound(a.eos, 1., en, pr, cs2);
    const double hbar = 0.197326;   // common.f90:51
    const double eta = hbar * a.vp.eta_over_s * s;
    const double zeta = bulk_zeta(a.vp, eta, cs2, temp, s);
    const double g3 = a.coordinates == BJORKEN ? tau0 * tau0 : 1.;
    V[KPIZZ * a.N + c] = -g3 * 4. * eta / (3. * tau0);
    V[KPIXX * a.N + c] = 2. * eta / (3. * tau0);
    V[KPIYY * a.N + c] = 2. * eta / (3. * tau0);
    V[KPIBU * a.N + c] = -zeta / tau0;
  }
  if (err) atomicMax(status, err);
}

// ---------------------------------------------------------------------------
// Initial magnetic field of generate_B_field_profile (glaubermc.f90:
// 1009-1168): per cell, the field of the protons among the spectators and
// participants of nuclei A and B (participants weighted by pw), added to
// B^x, B^y, B^eta charge by charge in the reference's order (spectators A,
// participants A, spectators B, participants B).  The charges are staged in
// LDS; each list has already been cut at its first neutron.
// ---------------------------------------------------------------------------
struct BFieldArgs {
  double gam, vel, sel, sch, tmin;
  double w[4];   // prefactor per list: B_amplify_factor (times pw for participants)
  int n[4];      // protons per list
  int magfield_type;
};

constexpr int BFIELD_MAX = 2048;   // LDS: 32 KiB of (x, y) pairs

// one proton: the classical azimuthal field Bphi and the chiral fields Br,
// Bz of glaubermc.f90:1028-1056 (nucleus A, a = true) and 1099-1127 (B)
__device__ __forceinline__ void proton_field(bool a, double w, double xc, double yc, double zpos,
                                             const BFieldArgs& b, double tom, double tomz, double& bx,
                                             double& by, double& bz) {
  const double hbar = 0.197326, sqrh = 0.444214, qbig = 0.30282212;   // common.f90:51-55
  const double pi = 3.141592653589793;
  const double gam = b.gam, vel = b.vel, sel = b.sel, sch = b.sch;
  const double phi = atan2(yc, xc);
  const double xt2 = xc * xc + yc * yc;
  const double xt = sqrt(xt2);
  const double D = (gam * zpos) * (gam * zpos) + xt2;
  const double sqrD = sqrt(D);
  const double A = (sel * vel * gam / 2) * (gam * zpos - sqrD) / hbar;
  const double D15 = pow(D, 1.5);
  const double eA = exp(A);
  double bphi = 0., br = 0., bzz = 0.;
  if (b.magfield_type != 3)
    bphi = w * qbig / (4 * pi) * sqrh * (vel * gam * xt / D15) * (1 + sel * vel * gam * sqrD / (2 * hbar)) * eA;
  if (b.magfield_type != 2) {
    br = w * (-sch) * qbig / (8 * pi) * (vel * (gam * gam) * xt / D15) * (gam * zpos + A * sqrD) * eA / sqrh;
    bzz = w * sch * qbig / (8 * pi) * (vel * gam / D15) *
          ((gam * zpos) * (gam * zpos) * (1 + sel * vel * gam * sqrD / (2 * hbar)) +
           D * (1 - sel * vel * gam * sqrD / (2 * hbar))) *
          eA / sqrh;
  }
  double sp, cp;
  sincos(phi, &sp, &cp);
  if (a) {
    bx = bx + (bphi * (-sp) + br * cp) / tom;
    by = by + (bphi * cp + br * sp) / tom;
    bz = bz + bzz / tomz;
  } else {
    bx = bx + (bphi * sp + br * cp) / tom;
    by = by + (bphi * (-cp) + br * sp) / tom;
    bz = bz + (-bzz) / tomz;
  }
}

template <int SYS>
__global__ __launch_bounds__(UPDATE_BLOCK) void k_glauber_bfield(GridArgs a, BFieldArgs b, double x1, double dx,
                                                                 double y1, double dy, double z1, double dz,
                                                                 const double* __restrict__ xy,
                                                                 double* __restrict__ V) {
  __shared__ double2 s[BFIELD_MAX];
  const int ns = b.n[0] + b.n[1] + b.n[2] + b.n[3];
  for (int i = threadIdx.x; i < ns; i += UPDATE_BLOCK) s[i] = make_double2(xy[2 * i], xy[2 * i + 1]);
  __syncthreads();
  const long long c = (long long)blockIdx.x * UPDATE_BLOCK + threadIdx.x;
  if (c >= a.N) return;
  const long long N = a.N;
  const int ix = (int)(c % a.nx);
  const long long r = c / a.nx;
  const int iy = (int)(r % a.ny), iz = (int)(r / a.ny);
  const double x = x1 + ((double)(ix + 1) - 0.5) * dx;
  const double y = y1 + ((double)(iy + 1) - 0.5) * dy;
  const double z = z1 + ((double)(iz + 1) - 0.5) * dz;
  double tt, zcoord, tom, tomz;
  if (a.coordinates == MINKOWSKI) {
    zcoord = z; tt = b.tmin; tom = 1.; tomz = 1.;
  } else {
    tt = b.tmin * cosh(z); zcoord = b.tmin * sinh(z); tom = cosh(z); tomz = b.tmin;
  }
  constexpr int KB = SysT<SYS>::KB;   // B^x of the layout
  double bx = V[KB * N + c], by = V[(KB + 1) * N + c], bz = V[(KB + 2) * N + c];
  int i = 0;
  for (int l = 0; l < 4; l++) {
    const double zpos = l < 2 ? tt * b.vel + zcoord : tt * b.vel - zcoord;
    const int end = i + b.n[l];
    for (; i < end; i++) proton_field(l < 2, b.w[l], x - s[i].x, y - s[i].y, zpos, b, tom, tomz, bx, by, bz);
  }
  V[KB * N + c] = bx;
  V[(KB + 1) * N + c] = by;
  V[(KB + 2) * N + c] = bz;
}

// batched con2prim over n independent states (system_cons2prim)
template <int SYS, bool ZZ>
__global__ __launch_bounds__(UPDATE_BLOCK) void k_c2p_batch(GridArgs a, const double* __restrict__ Ub,
                                                            double* __restrict__ Vb,
                                                            int* __restrict__ status, long long n) {
  constexpr int NV = SysT<SYS>::NV;
  const long long i = (long long)blockIdx.x * UPDATE_BLOCK + threadIdx.x;
  if (i >= n) return;
  const Metric M = make_metric(a.coordinates, a.S->t);
  double u[NV], v[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) { u[k] = Ub[k * n + i]; v[k] = Vb[k * n + i]; }
  status[i] = cons2prim<SYS, ZZ>(a.eos, M, u, v, a.algo_solver);
#pragma unroll
  for (int k = 0; k < NV; k++) Vb[k * n + i] = v[k];
}

}  // namespace echo
