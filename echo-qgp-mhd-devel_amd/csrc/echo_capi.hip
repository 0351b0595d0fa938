// C-ABI of the MI355X grid update (declared in include/echo_amd.h): device
// buffers, launch sequence of one RK step, and the HIP graph that replays a
// whole step without host involvement.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <type_traits>
#include <algorithm>
#include <vector>

#include "../../include/echo_amd.h"
#include "grid.hpp"

using namespace echo;

// RK2/RK3 Shu-Osher coefficients crk(1:2, irk, nrk) of evolve.f90:122-125
static void rk_coeffs(int nrk, int irk, double& c1, double& c2) {
  static const double C[2][3][3] = {
      {{0., 0., 0.}, {0., 1. / 2., 3. / 4.}, {0., 0., 1. / 3.}},
      {{1., 1., 1.}, {0., 1. / 2., 1. / 4.}, {0., 0., 2. / 3.}}};
  c1 = C[0][irk - 1][nrk - 1];
  c2 = C[1][irk - 1][nrk - 1];
}


// ---------------------------------------------------------------------------
// kernel dispatch
// ---------------------------------------------------------------------------
// flux launch of one sweep with the grid's tile shape; the tiled variants
// are instantiated for the WENO5 and PPM4 reconstructions
template <int DIR, int REC, int SYS, bool ZZ>
static void launch_flux_rec(echo_grid* g, double* partial) {
  dim3 grid(g->nb[DIR - 1]), block(FLUX_BLOCK);
  double* F = g->F[DIR - 1];
  double* L = g->L[DIR - 1];
  if constexpr (DIR > 1 && (REC == REC_WENO5 || REC == REC_PPM4)) {
    if (g->tile[DIR - 1] == 2) {
      k_flux<DIR, REC, SYS, ZZ, 2><<<grid, block, 0, g->stream>>>(g->ga, g->V, F, L, partial);
      return;
    }
    if (g->tile[DIR - 1] == 4) {
      k_flux<DIR, REC, SYS, ZZ, 4><<<grid, block, 0, g->stream>>>(g->ga, g->V, F, L, partial);
      return;
    }
  }
  k_flux<DIR, REC, SYS, ZZ, 1><<<grid, block, 0, g->stream>>>(g->ga, g->V, F, L, partial);
}

template <int DIR, int SYS, bool ZZ>
static void launch_flux_dir(echo_grid* g, double* partial) {
  switch (g->cfg.recal) {
    case REC_TVD2: launch_flux_rec<DIR, REC_TVD2, SYS, ZZ>(g, partial); break;
    case REC_CENO3: launch_flux_rec<DIR, REC_CENO3, SYS, ZZ>(g, partial); break;
    case REC_WENO3: launch_flux_rec<DIR, REC_WENO3, SYS, ZZ>(g, partial); break;
    case REC_PPM4: launch_flux_rec<DIR, REC_PPM4, SYS, ZZ>(g, partial); break;
    case REC_MPE3: launch_flux_rec<DIR, REC_MPE3, SYS, ZZ>(g, partial); break;
    case REC_MPE5: launch_flux_rec<DIR, REC_MPE5, SYS, ZZ>(g, partial); break;
    default: launch_flux_rec<DIR, REC_WENO5, SYS, ZZ>(g, partial); break;
  }
}

// launch blocks of a sweep for tile shape ty (padded for xcd_block)
static unsigned flux_blocks(const echo_config& c, int d, int ty) {
  long long t = d == 0 ? flux_tiles<1>(c.nx, c.ny, c.nz, 1)
              : d == 1 ? flux_tiles<2>(c.nx, c.ny, c.nz, ty) : flux_tiles<3>(c.nx, c.ny, c.nz, ty);
  return (unsigned)((t + NXCD - 1) / NXCD * NXCD);
}

static bool tiles_supported(const echo_config& c) { return c.recal == REC_WENO5 || c.recal == REC_PPM4; }

// launch blocks of the cell updates (update_cell), padded for xcd_block
static unsigned update_blocks(const echo_grid* g) {
  if (g->ga.upd_tz == 1) return nblocks_xcd(g->N, UPDATE_BLOCK);
  const long long t = (long long)((g->cfg.nx + UPD_TX - 1) / UPD_TX) * g->cfg.ny * ((g->cfg.nz + UPD_TZ - 1) / UPD_TZ);
  return (unsigned)((t + NXCD - 1) / NXCD * NXCD);
}

template <int SYS, bool ZZ>
static void enqueue_stage_t(echo_grid* g, int irk) {
  const int nrk = g->cfg.nrk;
  double c1, c2;
  rk_coeffs(nrk, irk, c1, c2);
  double* part[3] = {g->partial, g->partial + g->nb[0], g->partial + g->nb[0] + g->nb[1]};
  if (g->timing) hipEventRecord(g->ev[0], g->stream);
  if (g->active[0]) launch_flux_dir<1, SYS, ZZ>(g, part[0]);
  if (g->timing) hipEventRecord(g->ev[1], g->stream);
  if (g->active[1]) launch_flux_dir<2, SYS, ZZ>(g, part[1]);
  if (g->timing) hipEventRecord(g->ev[2], g->stream);
  if (g->active[2]) launch_flux_dir<3, SYS, ZZ>(g, part[2]);
  if (g->timing) hipEventRecord(g->ev[3], g->stream);
  if (irk == 1) {
    k_cfl<<<1, 1024, 0, g->stream>>>(g->ga, g->partial, g->nb[0], g->nb[1], g->nb[2]);
  }
  if (g->timing) hipEventRecord(g->ev[4], g->stream);
  // stage 1 reads u^n from U[0] and writes U[1]; later stages update U[1]
  // in place; the last stage writes u^{n+1} back into U[0] (each cell reads
  // its own u0 before overwriting it), so buffer roles repeat every step.
  const bool last = (irk == nrk);
  double* Ucur = (irk == 1) ? g->U[0] : g->U[1];
  double* Uout = last ? g->U[0] : g->U[1];
  k_update<SYS, ZZ><<<update_blocks(g), UPDATE_BLOCK, 0, g->stream>>>(
      g->ga, g->V, g->U[0], Ucur, Uout, g->F[0], g->F[1], g->F[2], g->L[0], g->L[1], g->L[2],
      g->VS[0], g->VS[1], g->fix, c1, c2, irk == 1 ? 1 : 0, last ? 1 : 0);
  if (g->net) enqueue_net_c2p(g, Uout);   // ideal MHD with the learned inversion
  if (g->timing) hipEventRecord(g->ev[5], g->stream);
  // con2prim repair of the cells k_update could not invert (rare; returns at
  // once when there are none); it writes the next step's vnewstep too,
  // before k_end_step flips the parity
  k_fixup<SYS, ZZ><<<1, FIX_BLOCK, 0, g->stream>>>(g->ga, g->V, Uout, g->fix, g->VS[0],
                                                    g->VS[1], last ? 1 : 0, ImexView{nullptr, nullptr, 0}, 1);
}

// one stage of IMEX-SSP3(4,3,3) (kernels.hpp: k_imex_k1, k_imex_k2,
// k_imex_u2v, k_update_imex).  The launches are shifted against the scheme's
// stages: stage 1 sets dt and solves the implicit parts of K1 and K2 (the
// explicit rate of K1 has weight zero: on the resistive systems, which clean
// at the speed of light, it needs no sweep; on the others the sweep of K1
// gives maxval(aflux), the cleaning speed of stage 2, and K1 = K2 = u0 is
// inverted once); stages 2 and 3 sweep K2 (K3), form their rates and solve
// K3 (K4); stage 4 sweeps K4 and forms the step's result.  Each launch that
// inverts is followed by the repair of the cells it could not invert.
template <int SYS, bool ZZ>
static void enqueue_imex_stage_t(echo_grid* g, int irk) {
  if constexpr (!SysT<SYS>::RES) {
    double* part[3] = {g->partial, g->partial + g->nb[0], g->partial + g->nb[0] + g->nb[1]};
    double* Ust = g->U[1];
    auto fixup = [&](double* U, int last, int phase) {
      k_fixup<SYS, ZZ><<<1, FIX_BLOCK, 0, g->stream>>>(g->ga, g->V, U, g->fix, g->VS[0], g->VS[1], last,
                                                        ImexView{g->U[0], g->imex.P4, phase}, last);
    };
    if (g->timing) hipEventRecord(g->ev[0], g->stream);
    if (g->active[0]) launch_flux_dir<1, SYS, ZZ>(g, part[0]);
    if (g->timing) hipEventRecord(g->ev[1], g->stream);
    if (g->active[1]) launch_flux_dir<2, SYS, ZZ>(g, part[1]);
    if (g->timing) hipEventRecord(g->ev[2], g->stream);
    if (g->active[2]) launch_flux_dir<3, SYS, ZZ>(g, part[2]);
    if (g->timing) hipEventRecord(g->ev[3], g->stream);
    // dt from the speeds of the step's first sweeps, as in the RK steps
    if (irk == 1) k_cfl<<<1, 1024, 0, g->stream>>>(g->ga, g->partial, g->nb[0], g->nb[1], g->nb[2]);
    if (g->timing) hipEventRecord(g->ev[4], g->stream);
    if (irk == 1) {
      // K2 = u0 (evolve.f90:232-233): its inversion; the repair also sets
      // the cleaning speed of stage 2 from this stage's sweeps
      k_imex_u2v<SYS, ZZ><<<nblocks_xcd(g->N, UPDATE_BLOCK), UPDATE_BLOCK, 0, g->stream>>>(g->ga, g->V, g->U[0],
                                                                                          g->fix);
      if (g->timing) hipEventRecord(g->ev[5], g->stream);
      fixup(Ust, 0, 2);
      return;
    }
#define ECHO_IMEX_UPDATE(ST)                                                                        \
  k_update_imex<SYS, ZZ, ST><<<update_blocks(g), UPDATE_BLOCK, 0, g->stream>>>(                    \
      g->ga, g->V, g->U[0], Ust, g->F[0], g->F[1], g->F[2], g->L[0], g->L[1], g->L[2], g->VS[0], \
      g->VS[1], g->imex, g->fix)
    if (irk == 2) ECHO_IMEX_UPDATE(2);
    else if (irk == 3) ECHO_IMEX_UPDATE(3);
    else ECHO_IMEX_UPDATE(4);
#undef ECHO_IMEX_UPDATE
    if (g->timing) hipEventRecord(g->ev[5], g->stream);
    if (irk == 4) fixup(g->U[0], 1, 0);
    else fixup(Ust, 0, irk + 1);
  } else {
    const unsigned nbu = nblocks_xcd(g->N, UPDATE_BLOCK);
    double* part[3] = {g->partial, g->partial + g->nb[0], g->partial + g->nb[0] + g->nb[1]};
    double* Ust = g->U[1];
    // the repair after the implicit solve of K_phase (phase 0: the step's
    // result), reading U* through ImexView
    auto fixup = [&](double* U, int last, int phase) {
      k_fixup<SYS, ZZ><<<1, FIX_BLOCK, 0, g->stream>>>(g->ga, g->V, U, g->fix, g->VS[0], g->VS[1], last,
                                                        ImexView{g->U[0], g->imex.P4, phase}, last);
    };
    if (g->timing) hipEventRecord(g->ev[0], g->stream);
    if (irk == 1) {
      if (g->timing) for (int i = 1; i < 4; i++) hipEventRecord(g->ev[i], g->stream);
      k_cfl<<<1, 1024, 0, g->stream>>>(g->ga, g->partial, g->nb[0], g->nb[1], g->nb[2]);
      if (g->timing) hipEventRecord(g->ev[4], g->stream);
      k_imex_k1<SYS, ZZ><<<nbu, UPDATE_BLOCK, 0, g->stream>>>(g->ga, g->V, g->U[0], Ust, g->fix);
      fixup(Ust, 0, 1);
      k_imex_k2<SYS, ZZ><<<nbu, UPDATE_BLOCK, 0, g->stream>>>(g->ga, g->V, g->U[0], Ust, g->imex, g->fix);
      if (g->timing) hipEventRecord(g->ev[5], g->stream);
      fixup(Ust, 0, 2);
      return;
    }
    if (g->active[0]) launch_flux_dir<1, SYS, ZZ>(g, part[0]);
    if (g->timing) hipEventRecord(g->ev[1], g->stream);
    if (g->active[1]) launch_flux_dir<2, SYS, ZZ>(g, part[1]);
    if (g->timing) hipEventRecord(g->ev[2], g->stream);
    if (g->active[2]) launch_flux_dir<3, SYS, ZZ>(g, part[2]);
    if (g->timing) {
      hipEventRecord(g->ev[3], g->stream);
      hipEventRecord(g->ev[4], g->stream);
    }
#define ECHO_IMEX_UPDATE(ST)                                                                        \
  k_update_imex<SYS, ZZ, ST><<<update_blocks(g), UPDATE_BLOCK, 0, g->stream>>>(                    \
      g->ga, g->V, g->U[0], Ust, g->F[0], g->F[1], g->F[2], g->L[0], g->L[1], g->L[2], g->VS[0], \
      g->VS[1], g->imex, g->fix)
    if (irk == 2) ECHO_IMEX_UPDATE(2);
    else if (irk == 3) ECHO_IMEX_UPDATE(3);
    else ECHO_IMEX_UPDATE(4);
#undef ECHO_IMEX_UPDATE
    if (g->timing) hipEventRecord(g->ev[5], g->stream);
    if (irk == 4) fixup(g->U[0], 1, 0);
    else fixup(Ust, 0, irk + 1);
  }
}

static void enqueue_stage(echo_grid* g, int irk) {
  dispatch_sys(g->cfg, [&](auto sys, auto zz) {
    constexpr int S = decltype(sys)::value;
    constexpr bool Z = decltype(zz)::value;
    if (g->cfg.nrk == 4) enqueue_imex_stage_t<S, Z>(g, irk);
    else enqueue_stage_t<S, Z>(g, irk);
  });
}

// one flux sweep (direction d = 0, 1, 2) with the current tile shape
static void enqueue_flux(echo_grid* g, int d) {
  double* part[3] = {g->partial, g->partial + g->nb[0], g->partial + g->nb[0] + g->nb[1]};
  dispatch_sys(g->cfg, [&](auto sys, auto zz) {
    constexpr int S = decltype(sys)::value;
    constexpr bool Z = decltype(zz)::value;
    if (d == 0) launch_flux_dir<1, S, Z>(g, part[0]);
    else if (d == 1) launch_flux_dir<2, S, Z>(g, part[1]);
    else launch_flux_dir<3, S, Z>(g, part[2]);
  });
}

static int set_tile(echo_grid* g, int d, int ty) {
  if (d < 0 || d > 2) return fail(ECHO_E_ARG, "sweep index must be 0, 1 or 2");
  if (ty != 1 && ty != 2 && ty != 4) return fail(ECHO_E_ARG, "tile rows must be 1, 2 or 4");
  if (ty != 1 && (d == 0 || !tiles_supported(g->cfg)))
    return fail(ECHO_E_ARG, "tiled sweeps exist for y and z with WENO5 or PPM4");
  g->tile[d] = ty;
  g->nb[d] = (int)flux_blocks(g->cfg, d, ty);
  drop_graph(g);
  if (g->graph_dirty) *g->graph_dirty = true;
  return 0;
}

// autotuner: times each tile shape of the y and z sweeps on the current
// state (reps launches each, HIP events on the grid's stream) and keeps the
// fastest.  The sweeps only write the face buffers, so the state is unchanged.
static int tune_grid(echo_grid* g, int reps, double* ms_out) {
  for (int d = 1; d < 3; d++) {
    if (!g->active[d] || !tiles_supported(g->cfg)) continue;
    double best = 1e300;
    int best_ty = 1;
    for (int ty : {1, 2, 4}) {
      if (int e = set_tile(g, d, ty)) return e;
      enqueue_flux(g, d);   // warm-up
      HIPCHECK(hipEventRecord(g->ev[6], g->stream));
      for (int r = 0; r < reps; r++) enqueue_flux(g, d);
      HIPCHECK(hipEventRecord(g->ev[7], g->stream));
      HIPCHECK(hipEventSynchronize(g->ev[7]));
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, g->ev[6], g->ev[7]));
      const double per = ms / reps;
      if (ms_out) ms_out[3 * d + (ty == 1 ? 0 : ty == 2 ? 1 : 2)] = per;
      if (per < best) { best = per; best_ty = ty; }
    }
    if (int e = set_tile(g, d, best_ty)) return e;
  }
  HIPCHECK(hipGetLastError());
  return 0;
}

static void enqueue_end_step(echo_grid* g) {
  k_end_step<<<1, 64, 0, g->stream>>>(g->S, g->cfg.coordinates, g->cfg.viscosity ? 1 : 0);
}

static void enqueue_check(echo_grid* g) {
  if (!g->stop_check) return;
  const unsigned nb = nblocks(g->N, CHECK_BLOCK);
  k_check_temp<<<nb, CHECK_BLOCK, 0, g->stream>>>(g->ga, g->V, g->partial);
  k_check_decide<<<1, 1024, 0, g->stream>>>(g->S, g->partial, (int)nb);
}

// krh..kvold_end of vold / uold (common.f90:379-386): up to B^eta with MHD
// (after the viscous variables on the viscous resistive grid), else up to p
static int dump_kend(const echo_grid* g) {
  return g->cfg.mhd ? (g->cfg.mhd == 2 && g->cfg.viscosity ? 15 : 8) : 5;
}

// vnewstep / unewstep of the restart dump, at the start of every step
// (all three ways of stepping: captured steps, timed steps, single stages)
static void enqueue_dump_copy(echo_grid* g) {
  if (!g->dump_track) return;
  const long long n = (long long)dump_kend(g) * g->N;
  k_dump_copy<<<nblocks(n, UPDATE_BLOCK), UPDATE_BLOCK, 0, g->stream>>>(g->S, g->V, g->U[0], g->dump_buf[0],
                                                                        g->dump_buf[1], n);
}

static void enqueue_step(echo_grid* g) {
  enqueue_check(g);
  enqueue_dump_copy(g);
  for (int irk = 1; irk <= g->cfg.nrk; irk++) enqueue_stage(g, irk);
  enqueue_end_step(g);
}

static void release(echo_grid* g) {
  if (!g) return;
  hipSetDevice(g->dev);
  if (g->writer) {
    g->writer->join();
    delete g->writer;
    g->writer = nullptr;
  }
  if (g->frame_dev) hipFree(g->frame_dev);
  if (g->frame_host) hipHostFree(g->frame_host);
  if (g->frame_ready) hipEventDestroy(g->frame_ready);
  if (g->frame_copied) hipEventDestroy(g->frame_copied);
  if (g->copy_stream) hipStreamDestroy(g->copy_stream);
  if (g->gexec) hipGraphExecDestroy(g->gexec);
  if (g->graph) hipGraphDestroy(g->graph);
  hipFree(g->V); hipFree(g->U[0]); hipFree(g->U[1]);
  if (g->owns_scratch) {
    for (int d = 0; d < 3; d++) { hipFree(g->F[d]); hipFree(g->L[d]); }
    hipFree(g->partial);
  }
  hipFree(g->VS[0]); hipFree(g->VS[1]); hipFree(g->tab); hipFree(g->tabix);
  hipFree(g->imex.P4); hipFree(g->imex.ACC);
  if (g->owns_scratch) {
    hipFree(g->fix.Pold); hipFree(g->fix.Upre); hipFree(g->fix.failbits);
    hipFree(g->fix.list); hipFree(g->fix.level);
  }
  hipFree(g->S); hipFree(g->status);
  hipFree(g->fo_f[0]); hipFree(g->fo_f[1]); hipFree(g->fo_f[2]); hipFree(g->fo_vb); hipFree(g->fo_mask);
  hipFree(g->fo_bsum); hipFree(g->fo_status); hipFree(g->fo_rec);
  hipFree(g->dump_buf[0]); hipFree(g->dump_buf[1]);
  hipFree(g->net_scratch);
  for (auto& e : g->ev) if (e) hipEventDestroy(e);
  if (g->stream && g->owns_stream) hipStreamDestroy(g->stream);
  delete g;
}

extern "C" {

const char* echo_last_error(void) { return echo::g_err.c_str(); }

// error text of the host-only translation units (glauber_mc.cpp)
__attribute__((visibility("hidden"))) void echo_set_error_text(const char* msg) { echo::g_err = msg; }

}  // extern "C"

// words of the failure bitmap of a grid of n cells
static long long fix_words(long long n) { return (n + 31) / 32; }

// components per face of a flux buffer: nv, and for resistive MHD the
// difference of the normal E across the face (for div E)
static int face_slots(const echo_config& c, int nv) { return c.mhd == 2 ? nv + 1 : nv; }

// Set-up of one grid.  stream: launch stream to use (not owned) or null for
// a stream of its own; share_scratch: leave F, L and partial unallocated
// (an ensemble points them at its shared pool).
static int grid_create(const echo_config* cfg, const double* eos_rows, int32_t n_rows, int32_t device,
                       hipStream_t stream, bool share_scratch, echo_grid** out) {
  if (!cfg || !out) return fail(ECHO_E_ARG, "echo_create: null argument");
  *out = nullptr;
  const echo_config& c = *cfg;
  if (c.nx < 1 || c.ny < 1 || c.nz < 1) return fail(ECHO_E_ARG, "grid sizes must be >= 1");
  if (c.mhd != 0 && c.mhd != 1 && c.mhd != 2) return fail(ECHO_E_ARG, "MHD must be 0, 1 or 2");
  if (c.mhd == 1 && c.viscosity)
    return fail(ECHO_E_ARG, "viscosity with ideal MHD is not supported (the resistive system, MHD=2, "
                            "takes viscosity)");
  if (c.mhd == 2 && !c.divclean)
    return fail(ECHO_E_ARG, "resistive MHD (MHD=2) needs divergence cleaning (DIVCLEAN=1)");
  if (c.mhd == 2 && (c.sigma_el < 0. || c.sigma_chi < 0.))
    return fail(ECHO_E_ARG, "SIGMA_EL and SIGMA_CHI must be >= 0");
  if (c.mhd == 1 && !((c.algo_solver == 4 && c.eqstate == 2) || (c.algo_solver == 5 && c.eqstate == 1) ||
                 (c.algo_solver == 6 && c.eqstate == 1 && c.coordinates == 1) ||
                 ((c.algo_solver == 1 || c.algo_solver == 2) && (c.eqstate == 1 || c.eqstate == 3))))
    return fail(ECHO_E_ARG, "MHD needs ALG_SOLV 1 or 2 with EQ_STATE 1 or 3, 4 with EQ_STATE 2, "
                            "or 5 (6 in Minkowski) with EQ_STATE 1 (got %d, %d)", c.algo_solver, c.eqstate);
  if (c.coordinates != 1 && c.coordinates != 2) return fail(ECHO_E_ARG, "COORD must be 1 or 2");
  if (c.derivatives_out)
    return fail(ECHO_E_ARG, "derivativ = 1: the derivative output of out_derivatives (out.f90:860-1250) "
                            "is not built");
  if (c.nrk == 4 && !c.mhd)
    return fail(ECHO_E_ARG, "INT_ALGO 4 (IMEX-SSP3(4,3,3)) is for MHD grids (common.f90:2076-2083)");
  if (c.nrk != 2 && c.nrk != 3 && c.nrk != 4)
    return fail(ECHO_E_ARG, "INT_ALGO must be RK2 (2), RK3 (3) or IMEX-SSP3(4,3,3) (4)");
  if (c.boundary_cond != 0 && c.boundary_cond != 2)
    return fail(ECHO_E_ARG, "boundary_cond must be 0 (periodic) or 2 (outflow)");
  if (c.recal < 1 || c.recal > 7) return fail(ECHO_E_ARG, "unsupported reconstruction code %d", c.recal);
  if (c.eqstate < 1 || c.eqstate > 3) return fail(ECHO_E_ARG, "EQ_STATE must be 1, 2 or 3");
  if (c.eqstate == 3 && (!eos_rows || n_rows < 2)) return fail(ECHO_E_ARG, "tabulated EoS needs rows");
  for (int d = 0; d < 3; d++) {
    int n = d == 0 ? c.nx : d == 1 ? c.ny : c.nz;
    if (n > 1 && n < 3) return fail(ECHO_E_ARG, "active dimensions need at least 3 cells");
  }
  echo_grid* g = new echo_grid();
  g->cfg = c;
  g->dev = device;
  hipError_t he = hipSetDevice(device);
  if (he != hipSuccess) { delete g; return fail(ECHO_E_HIP, "hipSetDevice: %s", hipGetErrorString(he)); }
  g->nv = c.mhd == 2 ? (c.viscosity ? 19 : 12) : c.mhd ? (c.divclean ? 9 : 8) : (c.viscosity ? 12 : 5);
  g->N = (long long)c.nx * c.ny * c.nz;
  g->nf[0] = (long long)(c.nx + 1) * c.ny * c.nz;
  g->nf[1] = (long long)c.nx * (c.ny + 1) * c.nz;
  g->nf[2] = (long long)c.nx * c.ny * (c.nz + 1);
  g->active[0] = c.nx > 1; g->active[1] = c.ny > 1; g->active[2] = c.nz > 1;
  long long nb_cap = 1 + (long long)nblocks(g->N, CHECK_BLOCK);
  // z sweeps start on two-row tiles where they exist: with one row the
  // stencil planes are fetched about six times over (measured FETCH_SIZE)
  if (tiles_supported(c) && c.nz > 1) g->tile[2] = 2;
  for (int d = 0; d < 3; d++) {
    g->nb[d] = (int)flux_blocks(c, d, g->tile[d]);
    unsigned mx = 0;
    for (int ty : {1, 2, 4}) mx = std::max(mx, flux_blocks(c, d, ty));
    nb_cap += mx;
  }
  const size_t szs = sizeof(double) * (size_t)g->nv * g->N;
#define ALLOC(p, bytes)                                                           \
  do {                                                                            \
    if (hipMalloc(&(p), (bytes)) != hipSuccess) {                                 \
      release(g);                                                                 \
      return fail(ECHO_E_NOMEM, "hipMalloc of %zu bytes failed", (size_t)(bytes)); \
    }                                                                             \
  } while (0)
  g->owns_scratch = !share_scratch;
  g->partial_cap = nb_cap;
  ALLOC(g->V, szs);
  ALLOC(g->U[0], szs);
  ALLOC(g->U[1], szs);
  for (int d = 0; d < 3 && !share_scratch; d++) {
    if (!g->active[d]) continue;
    ALLOC(g->F[d], sizeof(double) * (size_t)face_slots(c, g->nv) * g->nf[d]);
    if (c.viscosity) ALLOC(g->L[d], sizeof(double) * 4 * (size_t)g->nf[d]);
  }
  if (c.viscosity) {
    ALLOC(g->VS[0], sizeof(double) * 3 * (size_t)g->N);
    ALLOC(g->VS[1], sizeof(double) * 3 * (size_t)g->N);
  }
  if (!share_scratch) ALLOC(g->partial, sizeof(double) * (size_t)nb_cap);
  ALLOC(g->S, sizeof(DevScalars));
  if (c.nrk == 4) {
    ALLOC(g->imex.P4, szs);
    ALLOC(g->imex.ACC, szs);
  }
  g->fix_upre_cap = (c.viscosity && c.smooth_temp > 0.) ? (long long)g->nv * g->N : 0;
  if (!share_scratch) {
    ALLOC(g->fix.Pold, sizeof(double) * (size_t)g->N);
    if (g->fix_upre_cap) ALLOC(g->fix.Upre, sizeof(double) * (size_t)g->fix_upre_cap);
    ALLOC(g->fix.failbits, sizeof(unsigned) * (size_t)fix_words(g->N));
    ALLOC(g->fix.list, sizeof(int) * (size_t)g->N);
    ALLOC(g->fix.level, sizeof(int) * (size_t)g->N);
  }
  ALLOC(g->status, sizeof(int));
  if (stream) {
    g->stream = stream;
    g->owns_stream = false;
  } else if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
    release(g);
    return fail(ECHO_E_HIP, "stream creation failed");
  }
  // EoS: the table conversion of read_data_from_tabulated_file (eos.f90:323-327)
  Eos E{};
  E.eqstate = c.eqstate;
  E.enezero = c.enezero;
  std::vector<double> tab;
  if (c.eqstate == 3) {
    const double hc3 = .00768351;
    g->n_tab = n_rows;
    tab.resize(4 * (size_t)n_rows);
    for (int i = 0; i < n_rows; i++) {
      double T = eos_rows[4 * i], e = eos_rows[4 * i + 1], p = eos_rows[4 * i + 2], s = eos_rows[4 * i + 3];
      tab[i] = T;
      tab[n_rows + i] = e / hc3 * pow(T, 4.0);
      tab[2 * n_rows + i] = p / hc3 * pow(T, 4.0);
      tab[3 * n_rows + i] = s;
    }
    for (int i = 1; i < n_rows; i++) {
      if (!(tab[n_rows + i] > tab[n_rows + i - 1]) || !(tab[2 * n_rows + i] > tab[2 * n_rows + i - 1])) {
        release(g);
        return fail(ECHO_E_ARG, "EoS table: energy density and pressure must increase strictly");
      }
    }
    ALLOC(g->tab, sizeof(double) * tab.size());
    if (hipMemcpy(g->tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice) != hipSuccess) {
      release(g);
      return fail(ECHO_E_HIP, "table upload failed");
    }
    E.n = n_rows;
    // bin starts of the bracket accelerator (phys.hpp, tab_bracket_fast)
    std::vector<int> ix[2];
    int base[2], nb[2];
    for (int w = 0; w < 2; w++) {
      const double* a = tab.data() + (w == 0 ? n_rows : 2 * n_rows);
      float f1 = (float)a[1], fl = (float)a[n_rows - 1];
      unsigned u1, ul;
      memcpy(&u1, &f1, 4);
      memcpy(&ul, &fl, 4);
      base[w] = (int)(u1 >> TAB_SHIFT) - 1;
      nb[w] = (int)(ul >> TAB_SHIFT) - base[w] + 1;
      ix[w].assign(nb[w] + 1, 0);
      // start[b] = #{j : bin(a[j]) < b}
      std::vector<int> cnt(nb[w] + 1, 0);
      for (int j = 0; j < n_rows; j++) cnt[tab_bin(a[j], base[w], nb[w]) + 1]++;
      for (int b = 1; b <= nb[w]; b++) ix[w][b] = ix[w][b - 1] + cnt[b];
    }
    ALLOC(g->tabix, sizeof(int) * (ix[0].size() + ix[1].size()));
    if (hipMemcpy(g->tabix, ix[0].data(), sizeof(int) * ix[0].size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->tabix + ix[0].size(), ix[1].data(), sizeof(int) * ix[1].size(), hipMemcpyHostToDevice) != hipSuccess) {
      release(g);
      return fail(ECHO_E_HIP, "table index upload failed");
    }
    E.base_e = base[0]; E.nb_e = nb[0]; E.base_p = base[1]; E.nb_p = nb[1];
  }
#undef ALLOC
  // find_przero (eos.f90:192-218), evaluated on the host table copy
  {
    Eos Eh = E;
    if (c.eqstate == 3) {
      Eh.tt = tab.data(); Eh.te = tab.data() + n_rows; Eh.tp = tab.data() + 2 * n_rows; Eh.tc = tab.data() + 3 * n_rows;
    }
    double pz, d1, d2;
    int e = eos_pressure(Eh, 0., c.enezero, pz, d1, d2);
    if (e || pz <= 0.) { release(g); return fail(ECHO_E_ARG, "ENEZERO gives no positive pressure"); }
    E.przero = pz;
    Eh.przero = pz;
    DevScalars S0{};
    S0.t = c.tmin; S0.timeold = c.tmin; S0.timeinterval = 0.; S0.g_cov_3_old = 0.;
    S0.dt = c.maxdt; S0.tout = c.tmin + c.dtout; S0.dtout = c.dtout; S0.tmax = c.tmax;
    // resistive systems clean at the speed of light
    S0.glm_ch_old = c.mhd == 2 ? 1. : 0.5;
    S0.glm_ch = S0.glm_ch_old;
    S0.crash_cell = 0x7fffffff;
    S0.smooth_pressure = (c.viscosity && c.smooth_temp > 0.) ? pressure_from_temperature(Eh, c.smooth_temp) : 0.;
    if (hipMemcpy(g->S, &S0, sizeof S0, hipMemcpyHostToDevice) != hipSuccess) {
      release(g);
      return fail(ECHO_E_HIP, "scalar upload failed");
    }
  }
  g->tab_host = std::move(tab);
  if (c.eqstate == 3) {
    E.tt = g->tab; E.te = g->tab + n_rows; E.tp = g->tab + 2 * n_rows; E.tc = g->tab + 3 * n_rows;
    E.ixe = g->tabix;
    E.ixp = g->tabix + (E.nb_e + 1);
  }
  GridArgs& a = g->ga;
  a.nx = c.nx; a.ny = c.ny; a.nz = c.nz; a.coordinates = c.coordinates; a.bc = c.boundary_cond;
  a.N = g->N;
  // common_grid, common.f90:441-481
  a.ddx = c.nx > 1 ? 1. / ((c.x2 - c.x1) / c.nx) : 0.;
  a.ddy = c.ny > 1 ? 1. / ((c.y2 - c.y1) / c.ny) : 0.;
  a.ddz = c.nz > 1 ? 1. / ((c.z2 - c.z1) / c.nz) : 0.;
  a.maxspeed = c.maxspeed; a.cfl = c.cfl; a.maxdt = c.maxdt;
  a.eos = E;
  a.vp.eta_over_s = c.eta_over_s; a.vp.tau_pi_coeff = c.tau_pi_coeff;
  a.vp.bulkvis = c.viscosity && c.bulk; a.vp.init_type = c.init_type;
  a.enable_smooth = c.viscosity && c.smooth_temp > 0.;
  a.algo_solver = c.algo_solver;
  a.divclean = c.mhd && c.divclean;
  a.glm_alpha = c.glm_alpha;
  // conductivities in fm^-1 (hbar c = 0.197326 GeV fm, common.f90:51)
  a.res = c.mhd == 2;
  a.sig = c.sigma_el / 0.197326;
  a.sch = c.sigma_chi / 0.197326;
  a.ext_c2p = 0;
  // z-row tiles for the MHD systems, whose updates read the most face data
  // (measured: viscous resistive step 2.1% faster, ideal MHD 1% faster,
  // viscous hydro 0.6% slower than cells in storage order)
  a.upd_tz = (c.mhd && c.nz >= UPD_TZ && c.nx >= UPD_TX) ? UPD_TZ : 1;
  a.S = g->S;
  if (!share_scratch && hipMemset(g->fix.failbits, 0, sizeof(unsigned) * (size_t)fix_words(g->N)) != hipSuccess) {
    release(g);
    return fail(ECHO_E_HIP, "memset failed");
  }
  for (auto& e : g->ev) hipEventCreate(&e);
  *out = g;
  return 0;
}


// ---------------------------------------------------------------------------
// ensemble of independent event grids on one device: one stream, one shared
// pool for the face buffers (F, L, partial; sized for the largest event),
// and one captured graph that advances every event by one time step
// ---------------------------------------------------------------------------
struct echo_ensemble {
  int dev = 0;
  hipStream_t stream = nullptr;
  std::vector<echo_grid*> g;
  double* F[3] = {nullptr, nullptr, nullptr};
  double* L[3] = {nullptr, nullptr, nullptr};
  double* partial = nullptr;
  FixBufs fix{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool dirty = true;
  hipEvent_t ev[2] = {};
  double last_ms = 0.0;
};

static void release_ensemble(echo_ensemble* e) {
  if (!e) return;
  hipSetDevice(e->dev);
  if (e->gexec) hipGraphExecDestroy(e->gexec);
  if (e->graph) hipGraphDestroy(e->graph);
  for (echo_grid* g : e->g) release(g);
  for (int d = 0; d < 3; d++) { hipFree(e->F[d]); hipFree(e->L[d]); }
  hipFree(e->partial);
  hipFree(e->fix.Pold); hipFree(e->fix.Upre); hipFree(e->fix.failbits);
  hipFree(e->fix.list); hipFree(e->fix.level);
  for (auto& v : e->ev) if (v) hipEventDestroy(v);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
}

extern "C" {

int echo_create(const echo_config* cfg, const double* eos_rows, int32_t n_rows, int32_t device,
                echo_grid** out) {
  return grid_create(cfg, eos_rows, n_rows, device, nullptr, false, out);
}

void echo_destroy(echo_grid* g) { release(g); }

int echo_nv(const echo_grid* g) { return g ? g->nv : 0; }

int echo_upload_state(echo_grid* g, const double* v, const double* u) {
  if (!g || !v || !u) return fail(ECHO_E_ARG, "null argument");
  HIPCHECK(hipSetDevice(g->dev));
  const size_t sz = sizeof(double) * (size_t)g->nv * g->N;
  HIPCHECK(hipMemcpyAsync(g->V, v, sz, hipMemcpyHostToDevice, g->stream));
  HIPCHECK(hipMemcpyAsync(g->U[0], u, sz, hipMemcpyHostToDevice, g->stream));
  if (g->cfg.viscosity) {
    // vnewstep of the first step: the uploaded velocities
    HIPCHECK(hipMemcpyAsync(g->VS[0], v + g->N * 1, sizeof(double) * 3 * g->N, hipMemcpyHostToDevice, g->stream));
    int par = 0;
    HIPCHECK(hipMemcpyAsync(&g->S->vs_parity, &par, sizeof(int), hipMemcpyHostToDevice, g->stream));
  }
  HIPCHECK(hipStreamSynchronize(g->stream));
  return 0;
}

int echo_download_state(echo_grid* g, double* v, double* u) {
  if (!g) return fail(ECHO_E_ARG, "null argument");
  HIPCHECK(hipSetDevice(g->dev));
  const size_t sz = sizeof(double) * (size_t)g->nv * g->N;
  if (v) HIPCHECK(hipMemcpyAsync(v, g->V, sz, hipMemcpyDeviceToHost, g->stream));
  if (u) HIPCHECK(hipMemcpyAsync(u, g->U[0], sz, hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  return 0;
}

static int run_init(echo_grid* g);

int echo_init_from_prim(echo_grid* g, const double* v) {
  if (!g || !v) return fail(ECHO_E_ARG, "null argument");
  HIPCHECK(hipSetDevice(g->dev));
  const size_t sz = sizeof(double) * (size_t)g->nv * g->N;
  HIPCHECK(hipMemcpyAsync(g->V, v, sz, hipMemcpyHostToDevice, g->stream));
  return run_init(g);
}

// primitives of the Glauber-MC initialisation in V (no prim2cons yet)
static int glauber_primitives(echo_grid* g, const echo_glauber_params* p, const double* part_a, int32_t n_a,
                              const double* part_b, int32_t n_b, const double* coll, int32_t n_coll) {
  if (!g || !p) return fail(ECHO_E_ARG, "null argument");
  if (n_a < 0 || n_b < 0 || n_coll < 0 || (n_a && !part_a) || (n_b && !part_b) || (n_coll && !coll))
    return fail(ECHO_E_ARG, "invalid source arrays");
  const int ns = n_a + n_b + n_coll;
  if (ns < 1 || ns > GLAUBER_MAX_SRC) return fail(ECHO_E_ARG, "1..%d sources supported, got %d", GLAUBER_MAX_SRC, ns);
  if (!(p->sig_mc > 0.) || !(p->ybeam > 0.)) return fail(ECHO_E_ARG, "need sig_mc > 0 and ybeam > 0");
  HIPCHECK(hipSetDevice(g->dev));
  // centre of mass of the sources, in the reference's summation order
  // (glaubermc.f90:806-821)
  double cmx_p = 0., cmy_p = 0., cmx_c = 0., cmy_c = 0.;
  for (int i = 0; i < n_a; i++) { cmx_p = cmx_p + part_a[2 * i]; cmy_p = cmy_p + part_a[2 * i + 1]; }
  for (int i = 0; i < n_b; i++) { cmx_p = cmx_p + part_b[2 * i]; cmy_p = cmy_p + part_b[2 * i + 1]; }
  for (int i = 0; i < n_coll; i++) { cmx_c = cmx_c + coll[2 * i]; cmy_c = cmy_c + coll[2 * i + 1]; }
  const double ah = p->ah;
  const double den = (1 - ah) * (n_a + n_b) + ah * n_coll;
  GlauberArgs q;
  q.kappa = p->kappa; q.ah = ah; q.s2 = p->sig_mc * p->sig_mc; q.ybeam = p->ybeam; q.deta = p->deta;
  q.sigeta = p->sigeta; q.rhocenter = p->rhocenter;
  q.cmx = ((1 - ah) * cmx_p + ah * cmx_c) / den;
  q.cmy = ((1 - ah) * cmy_p + ah * cmy_c) / den;
  q.na = n_a; q.nb = n_b; q.nc = n_coll;
  std::vector<double> src(2 * (size_t)ns);
  if (n_a) memcpy(src.data(), part_a, sizeof(double) * 2 * n_a);
  if (n_b) memcpy(src.data() + 2 * n_a, part_b, sizeof(double) * 2 * n_b);
  if (n_coll) memcpy(src.data() + 2 * (n_a + n_b), coll, sizeof(double) * 2 * n_coll);
  double *dsrc = nullptr, *dsum = nullptr;
  const long long n2 = (long long)g->cfg.nx * g->cfg.ny;
  if (hipMalloc(&dsrc, sizeof(double) * src.size()) != hipSuccess ||
      hipMalloc(&dsum, sizeof(double) * 3 * n2) != hipSuccess) {
    hipFree(dsrc);
    return fail(ECHO_E_NOMEM, "source buffers");
  }
  const echo_config& c = g->cfg;
  const double dx = (c.x2 - c.x1) / c.nx, dy = (c.y2 - c.y1) / c.ny, dz = (c.z2 - c.z1) / c.nz;
  int zero = 0;
  hipError_t e = hipMemcpyAsync(dsrc, src.data(), sizeof(double) * src.size(), hipMemcpyHostToDevice, g->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(g->status, &zero, sizeof(int), hipMemcpyHostToDevice, g->stream);
  if (e == hipSuccess) {
    k_glauber_sums<<<nblocks(n2, GLAUBER_BLOCK), GLAUBER_BLOCK, 0, g->stream>>>(g->ga, q, c.x1, dx, c.y1, dy,
                                                                                dsrc, dsum);
    dispatch_sys(g->cfg, [&](auto sys, auto) {
      k_glauber_prim<decltype(sys)::value><<<nblocks(g->N, UPDATE_BLOCK), UPDATE_BLOCK, 0, g->stream>>>(
          g->ga, q, c.z1, dz, c.tmin, c.enezero, acos(-1.), dsum, g->V, g->status);
    });
    e = hipGetLastError();
  }
  int st = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&st, g->status, sizeof(int), hipMemcpyDeviceToHost, g->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  hipFree(dsrc);
  hipFree(dsum);
  if (e != hipSuccess) return fail(ECHO_E_HIP, "glauber initial state: %s", hipGetErrorString(e));
  return st;
}

int echo_init_glauber(echo_grid* g, const echo_glauber_params* p, const double* part_a, int32_t n_a,
                      const double* part_b, int32_t n_b, const double* coll, int32_t n_coll) {
  const int st = glauber_primitives(g, p, part_a, n_a, part_b, n_b, coll, n_coll);
  if (st) return st;
  return run_init(g);
}

int echo_init_glauber_mhd(echo_grid* g, const echo_glauber_params* p, const double* part_a, int32_t n_a,
                          const double* part_b, int32_t n_b, const double* coll, int32_t n_coll,
                          const echo_bfield_params* bp, const echo_nucleons* lists) {
  if (!g || !p || !bp || !lists) return fail(ECHO_E_ARG, "null argument");
  if (!g->cfg.mhd) return fail(ECHO_E_ARG, "the grid has no magnetic field (mhd = 0)");
  if (!(bp->gamma_col > 1.)) return fail(ECHO_E_ARG, "need gamma_col > 1");
  if (bp->magfield_type < 1 || bp->magfield_type > 3) return fail(ECHO_E_ARG, "magfield_type must be 1, 2 or 3");
  // each list up to its first neutron, as glaubermc.f90:1026-1058 reads it
  std::vector<double> xy;
  int cnt[4];
  for (int l = 0; l < 4; l++) {
    const echo_nucleons& L = lists[l];
    if (L.n < 0 || (L.n && (!L.xy || !L.type))) return fail(ECHO_E_ARG, "invalid nucleon list %d", l);
    int k = 0;
    while (k < L.n && L.type[k] == 0) k++;
    cnt[l] = k;
    xy.insert(xy.end(), L.xy, L.xy + 2 * k);
  }
  const int ns = cnt[0] + cnt[1] + cnt[2] + cnt[3];
  if (ns > BFIELD_MAX) return fail(ECHO_E_ARG, "at most %d protons, got %d", BFIELD_MAX, ns);
  int st = glauber_primitives(g, p, part_a, n_a, part_b, n_b, coll, n_coll);
  if (st) return st;
  const echo_config& c = g->cfg;
  BFieldArgs b;
  b.gam = bp->gamma_col;
  b.vel = sqrt(1. - 1. / (bp->gamma_col * bp->gamma_col));
  b.sel = c.sigma_el;
  b.sch = c.sigma_chi;
  b.tmin = c.tmin;
  b.w[0] = bp->b_amplify; b.w[1] = bp->b_amplify * bp->pw;
  b.w[2] = bp->b_amplify; b.w[3] = bp->b_amplify * bp->pw;
  for (int l = 0; l < 4; l++) b.n[l] = cnt[l];
  b.magfield_type = bp->magfield_type;
  double* dxy = nullptr;
  if (ns > 0 && hipMalloc(&dxy, sizeof(double) * xy.size()) != hipSuccess)
    return fail(ECHO_E_NOMEM, "proton buffer");
  hipError_t e = hipSuccess;
  if (ns > 0) e = hipMemcpyAsync(dxy, xy.data(), sizeof(double) * xy.size(), hipMemcpyHostToDevice, g->stream);
  if (e == hipSuccess) {
    const double dx = (c.x2 - c.x1) / c.nx, dy = (c.y2 - c.y1) / c.ny, dz = (c.z2 - c.z1) / c.nz;
    dispatch_sys(c, [&](auto sys, auto) {
      constexpr int SYS = decltype(sys)::value;
      // resistive grids: B at the layout's slots, E and psi start at zero
      // (init.f90:2732-2832 fills kbx..kbz only)
      if constexpr (SysT<SYS>::MHD || SysT<SYS>::RES)
        k_glauber_bfield<SYS><<<nblocks(g->N, UPDATE_BLOCK), UPDATE_BLOCK, 0, g->stream>>>(
            g->ga, b, c.x1, dx, c.y1, dy, c.z1, dz, dxy, g->V);
    });
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  if (dxy) hipFree(dxy);
  if (e != hipSuccess) return fail(ECHO_E_HIP, "initial magnetic field: %s", hipGetErrorString(e));
  return run_init(g);
}

// prim2cons (+ con2prim for non-viscous systems) of the primitives in V,
// as the reference initialisers do (init.f90:2937-2958)
static int run_init(echo_grid* g) {
  int zero = 0;
  HIPCHECK(hipMemcpyAsync(g->status, &zero, sizeof(int), hipMemcpyHostToDevice, g->stream));
  HIPCHECK(hipMemcpyAsync(&g->S->vs_parity, &zero, sizeof(int), hipMemcpyHostToDevice, g->stream));
  dim3 grid(nblocks(g->N, UPDATE_BLOCK));
  dispatch_sys(g->cfg, [&](auto sys, auto zz) {
    k_init<decltype(sys)::value, decltype(zz)::value>
        <<<grid, UPDATE_BLOCK, 0, g->stream>>>(g->ga, g->V, g->U[0], g->VS[0], g->status);
  });
  HIPCHECK(hipGetLastError());
  int st = 0;
  HIPCHECK(hipMemcpyAsync(&st, g->status, sizeof(int), hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  return st;
}

// the run's crash state, with a message saying what crashed it
static int read_crash(echo_grid* g) {
  DevScalars S;
  HIPCHECK(hipMemcpyAsync(&S, g->S, sizeof S, hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  if (!S.run_crashed) return 0;
  const int code = S.errcode > 0 ? S.errcode : ECHO_E_CRASHED;
  const int nx = g->cfg.nx, ny = g->cfg.ny;
  switch (S.crash_kind) {
    case CRASH_REPAIR: {
      const long long c = S.crash_cell;
      return fail(code, "cons2prim failed and its neighbour-average repair failed too (evolve_u2v, "
                        "evolve.f90:836-840) at cell ix=%lld iy=%lld iz=%lld (1-based), t=%.17g",
                  c % nx + 1, (c / nx) % ny + 1, c / ((long long)nx * ny) + 1, S.t);
    }
    case CRASH_FLUX:
      return fail(code, "flux computation failed with error code %d (system_flux, system.f90:1316-1583), "
                        "t=%.17g", S.errcode, S.t);
    case CRASH_SOURCE:
      return fail(code, "source terms failed with error code %d (system_source, system.f90:1587-1842), "
                        "t=%.17g", S.errcode, S.t);
    default:
      return fail(code, "the run has crashed (error code %d), t=%.17g", S.errcode, S.t);
  }
}

int echo_rk_stage(echo_grid* g, int32_t irk) {
  if (!g) return fail(ECHO_E_ARG, "null argument");
  if (irk != g->next_stage || irk < 1 || irk > g->cfg.nrk)
    return fail(ECHO_E_ARG, "stages must be called in order 1..nrk (expected %d)", g->next_stage);
  HIPCHECK(hipSetDevice(g->dev));
  if (irk == 1) {
    enqueue_check(g);
    enqueue_dump_copy(g);
  }
  enqueue_stage(g, irk);
  HIPCHECK(hipGetLastError());
  g->next_stage = irk + 1;
  return read_crash(g);
}

int echo_end_step(echo_grid* g) {
  if (!g) return fail(ECHO_E_ARG, "null argument");
  if (g->next_stage != g->cfg.nrk + 1) return fail(ECHO_E_ARG, "echo_end_step before the last stage");
  HIPCHECK(hipSetDevice(g->dev));
  enqueue_end_step(g);
  HIPCHECK(hipGetLastError());
  g->next_stage = 1;
  return read_crash(g);
}

int echo_steps(echo_grid* g, int32_t nsteps) {
  if (!g) return fail(ECHO_E_ARG, "null argument");
  if (g->next_stage != 1) return fail(ECHO_E_ARG, "echo_steps inside a step");
  HIPCHECK(hipSetDevice(g->dev));
  if (g->timing) {
    // timed path: plain launches with events around each kernel
    // means per launch over the stages that sweep (every RK stage; IMEX
    // stages 2-4, so the first IMEX stage's implicit solves are in no mean)
    double acc[4] = {0, 0, 0, 0};
    int nsw = 0;
    for (int s = 0; s < nsteps; s++) {
      enqueue_check(g);
      enqueue_dump_copy(g);
      for (int irk = 1; irk <= g->cfg.nrk; irk++) {
        enqueue_stage(g, irk);
        HIPCHECK(hipEventSynchronize(g->ev[5]));
        float ms;
        if (!(g->cfg.nrk == 4 && irk == 1)) {
          for (int d = 0; d < 3; d++) {
            hipEventElapsedTime(&ms, g->ev[d], g->ev[d + 1]);
            acc[d] += ms;
          }
          hipEventElapsedTime(&ms, g->ev[4], g->ev[5]);
          acc[3] += ms;
          nsw++;
        }
      }
      enqueue_end_step(g);
    }
    for (int i = 0; i < 4; i++) g->tms[i] = nsw ? acc[i] / nsw : 0.;
    HIPCHECK(hipGetLastError());
    return read_crash(g);
  }
  if (!g->gexec) {
    HIPCHECK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(g);
    hipGraph_t graph;
    HIPCHECK(hipStreamEndCapture(g->stream, &graph));
    g->graph = graph;
    HIPCHECK(hipGraphInstantiate(&g->gexec, graph, nullptr, nullptr, 0));
  }
  for (int s = 0; s < nsteps; s++) HIPCHECK(hipGraphLaunch(g->gexec, g->stream));
  HIPCHECK(hipGetLastError());
  return read_crash(g);
}

int echo_get_scalars(echo_grid* g, double* s) {
  if (!g || !s) return fail(ECHO_E_ARG, "null argument");
  HIPCHECK(hipSetDevice(g->dev));
  DevScalars S;
  HIPCHECK(hipMemcpyAsync(&S, g->S, sizeof S, hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  s[0] = S.t; s[1] = S.dt; s[2] = S.timeold; s[3] = S.timeinterval; s[4] = S.g_cov_3_old;
  s[5] = S.glm_ch_old; s[6] = S.tout; s[7] = S.iter; s[8] = S.run_crashed; s[9] = S.errcode;
  return 0;
}

int echo_get_repairs(echo_grid* g, int64_t* n) {
  if (!g || !n) return fail(ECHO_E_ARG, "null argument");
  HIPCHECK(hipSetDevice(g->dev));
  DevScalars S;
  HIPCHECK(hipMemcpyAsync(&S, g->S, sizeof S, hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  *n = S.nrepair;
  return 0;
}

int echo_set_bulk_peak(echo_grid* g, double zs_max, double t_c, double width) {
  if (!g) return fail(ECHO_E_ARG, "null argument");
  if (!g->cfg.viscosity || !g->cfg.bulk) return fail(ECHO_E_ARG, "echo_set_bulk_peak needs VISCOUS and BULK");
  if (!(zs_max >= 0.) || !(t_c > 0.) || !(width > 0.))
    return fail(ECHO_E_ARG, "echo_set_bulk_peak: need zs_max >= 0, t_c > 0, width > 0");
  if (g->next_stage != 1) return fail(ECHO_E_ARG, "echo_set_bulk_peak inside a step");
  g->ga.vp.zeta_peak = 1;
  g->ga.vp.zs_max = zs_max;
  g->ga.vp.zs_tc = t_c;
  g->ga.vp.zs_w = width;
  drop_graph(g);   // the captured launches hold the old parameters
  if (g->graph_dirty) *g->graph_dirty = true;
  return 0;
}

int echo_set_tout(echo_grid* g, double tout) {
  if (!g) return fail(ECHO_E_ARG, "null argument");
  HIPCHECK(hipSetDevice(g->dev));
  HIPCHECK(hipMemcpyAsync(&g->S->tout, &tout, sizeof(double), hipMemcpyHostToDevice, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  return 0;
}

int echo_cons2prim_batch(echo_grid* g, const double* u, double* v, int32_t* status, int64_t n) {
  if (!g || !u || !v || !status || n < 0) return fail(ECHO_E_ARG, "invalid argument");
  if (n == 0) return 0;
  HIPCHECK(hipSetDevice(g->dev));
  const size_t sz = sizeof(double) * (size_t)g->nv * n;
  double *du = nullptr, *dv = nullptr;
  int* ds = nullptr;
  HIPCHECK(hipMalloc(&du, sz));
  HIPCHECK(hipMalloc(&dv, sz));
  HIPCHECK(hipMalloc(&ds, sizeof(int) * n));
  HIPCHECK(hipMemcpyAsync(du, u, sz, hipMemcpyHostToDevice, g->stream));
  HIPCHECK(hipMemcpyAsync(dv, v, sz, hipMemcpyHostToDevice, g->stream));
  dim3 grid(nblocks(n, UPDATE_BLOCK));
  dispatch_sys(g->cfg, [&](auto sys, auto zz) {
    k_c2p_batch<decltype(sys)::value, decltype(zz)::value>
        <<<grid, UPDATE_BLOCK, 0, g->stream>>>(g->ga, du, dv, ds, n);
  });
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(v, dv, sz, hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipMemcpyAsync(status, ds, sizeof(int) * n, hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  hipFree(du); hipFree(dv); hipFree(ds);
  return 0;
}

int echo_set_stop(echo_grid* g, double temp_end, int32_t freeze_type) {
  if (!g) return fail(ECHO_E_ARG, "null argument");
  if (freeze_type != 0 && freeze_type != 1) return fail(ECHO_E_ARG, "freeze_type must be 0 or 1");
  if (g->next_stage != 1) return fail(ECHO_E_ARG, "echo_set_stop inside a step");
  HIPCHECK(hipSetDevice(g->dev));
  int zero = 0;
  HIPCHECK(hipMemcpyAsync(&g->S->temp_end, &temp_end, sizeof(double), hipMemcpyHostToDevice, g->stream));
  HIPCHECK(hipMemcpyAsync(&g->S->freeze_type, &freeze_type, sizeof(int), hipMemcpyHostToDevice, g->stream));
  HIPCHECK(hipMemcpyAsync(&g->S->temp_check, &zero, sizeof(int), hipMemcpyHostToDevice, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  g->stop_check = temp_end > 0.;
  drop_graph(g);
  if (g->graph_dirty) *g->graph_dirty = true;
  return 0;
}

int echo_set_dump(echo_grid* g, int32_t on) {
  if (!g) return fail(ECHO_E_ARG, "null argument");
  if (g->next_stage != 1) return fail(ECHO_E_ARG, "echo_set_dump inside a step");
  HIPCHECK(hipSetDevice(g->dev));
  if (on && !g->dump_buf[0]) {
    const size_t sz = sizeof(double) * 2 * dump_kend(g) * (size_t)g->N;
    for (int b = 0; b < 2; b++)
      if (hipMalloc(&g->dump_buf[b], sz) != hipSuccess) {
        hipFree(g->dump_buf[0]);
        g->dump_buf[0] = g->dump_buf[1] = nullptr;
        return fail(ECHO_E_NOMEM, "echo_set_dump: allocation of %zu bytes failed", 2 * sz);
      }
  }
  if (on) {
    // until a step has run, vold / uold are the current state (init.f90:3271-3272)
    DevScalars S;
    HIPCHECK(hipMemcpyAsync(&S, g->S, sizeof S, hipMemcpyDeviceToHost, g->stream));
    HIPCHECK(hipStreamSynchronize(g->stream));
    const size_t n = (size_t)dump_kend(g) * g->N;
    double* B = g->dump_buf[(S.iter - 1) & 1];
    HIPCHECK(hipMemcpyAsync(B, g->V, sizeof(double) * n, hipMemcpyDeviceToDevice, g->stream));
    HIPCHECK(hipMemcpyAsync(B + n, g->U[0], sizeof(double) * n, hipMemcpyDeviceToDevice, g->stream));
    HIPCHECK(hipStreamSynchronize(g->stream));
  }
  g->dump_track = on != 0;
  drop_graph(g);
  if (g->graph_dirty) *g->graph_dirty = true;
  return 0;
}

// the freeze-out part of the dump: thyp, time_arr, the slice the next
// echo_fo_slice compares against, field(2,:,:,:)
static int fo_dump_io(echo_grid* g, FILE* f, bool write, double* thyp) {
  const size_t N = (size_t)g->N, nvN = N * g->nv;
  std::vector<double> a(nvN), fld(N);
  double ta[2] = {g->fo_t1, g->fo_t2};
  if (write) {
    HIPCHECK(hipMemcpyAsync(a.data(), g->fo_vb, sizeof(double) * nvN, hipMemcpyDeviceToHost, g->stream));
    HIPCHECK(hipMemcpyAsync(fld.data(), g->fo_f[1], sizeof(double) * N, hipMemcpyDeviceToHost, g->stream));
    HIPCHECK(hipStreamSynchronize(g->stream));
    bool bad = fwrite(thyp, sizeof(double), 1, f) != 1;
    bad |= fwrite(ta, sizeof(double), 2, f) != 2;
    bad |= fwrite(a.data(), sizeof(double), nvN, f) != nvN;
    bad |= fwrite(fld.data(), sizeof(double), N, f) != N;
    return bad ? fail(ECHO_E_ARG, "dump: write failed") : 0;
  }
  bool bad = fread(thyp, sizeof(double), 1, f) != 1;
  bad |= fread(ta, sizeof(double), 2, f) != 2;
  bad |= fread(a.data(), sizeof(double), nvN, f) != nvN;
  bad |= fread(fld.data(), sizeof(double), N, f) != N;
  if (bad) return fail(ECHO_E_ARG, "restart: the file ends early");
  HIPCHECK(hipMemcpyAsync(g->fo_vb, a.data(), sizeof(double) * nvN, hipMemcpyHostToDevice, g->stream));
  HIPCHECK(hipMemcpyAsync(g->fo_f[1], fld.data(), sizeof(double) * N, hipMemcpyHostToDevice, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  g->fo_t1 = ta[0];
  g->fo_t2 = ta[1];
  g->fo_n = 0;
  g->fo_searched = false;
  return 0;
}

int echo_dump(echo_grid* g, const char* path, int32_t iout, double thyp) {
  if (!g || !path) return fail(ECHO_E_ARG, "null argument");
  if (!g->dump_track) return fail(ECHO_E_ARG, "echo_dump: call echo_set_dump(g, 1) before the steps to dump");
  if (g->next_stage != 1) return fail(ECHO_E_ARG, "echo_dump inside a step");
  HIPCHECK(hipSetDevice(g->dev));
  const size_t N = (size_t)g->N, nvN = N * g->nv, kN = N * dump_kend(g);
  DevScalars S;
  std::vector<double> v(nvN), u(nvN), old(2 * kN);
  HIPCHECK(hipMemcpyAsync(&S, g->S, sizeof S, hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  HIPCHECK(hipMemcpyAsync(v.data(), g->V, sizeof(double) * nvN, hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipMemcpyAsync(u.data(), g->U[0], sizeof(double) * nvN, hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipMemcpyAsync(old.data(), g->dump_buf[(S.iter - 1) & 1], sizeof(double) * 2 * kN,
                          hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  FILE* f = fopen(path, "wb");
  if (!f) return fail(ECHO_E_ARG, "echo_dump: cannot open %s", path);
  bool bad = fwrite(&iout, 4, 1, f) != 1;
  bad |= fwrite(&S.t, sizeof(double), 1, f) != 1;
  bad |= fwrite(&S.timeold, sizeof(double), 1, f) != 1;
  bad |= fwrite(&S.timeinterval, sizeof(double), 1, f) != 1;
  bad |= fwrite(v.data(), sizeof(double), nvN, f) != nvN;
  bad |= fwrite(u.data(), sizeof(double), nvN, f) != nvN;
  bad |= fwrite(old.data(), sizeof(double), 2 * kN, f) != 2 * kN;   // vold, then uold
  bad |= fwrite(&S.g_cov_3_old, sizeof(double), 1, f) != 1;
  int e = 0;
  if (!bad && g->fo_on) e = fo_dump_io(g, f, true, &thyp);
  bad |= fclose(f) != 0;
  if (e) return e;
  return bad ? fail(ECHO_E_ARG, "echo_dump: write to %s failed", path) : 0;
}

int echo_restart(echo_grid* g, const char* path, int32_t* iout, double* thyp) {
  if (!g || !path || !iout) return fail(ECHO_E_ARG, "null argument");
  if (g->next_stage != 1) return fail(ECHO_E_ARG, "echo_restart inside a step");
  HIPCHECK(hipSetDevice(g->dev));
  const size_t N = (size_t)g->N, nvN = N * g->nv, kN = N * dump_kend(g);
  FILE* f = fopen(path, "rb");
  if (!f) return fail(ECHO_E_ARG, "echo_restart: cannot open %s", path);
  int32_t io = 0;
  double t, timeold, timeinterval, g3old;
  std::vector<double> v(nvN), u(nvN), old(2 * kN);
  bool bad = fread(&io, 4, 1, f) != 1;
  bad |= fread(&t, sizeof(double), 1, f) != 1;
  bad |= fread(&timeold, sizeof(double), 1, f) != 1;
  bad |= fread(&timeinterval, sizeof(double), 1, f) != 1;
  bad |= fread(v.data(), sizeof(double), nvN, f) != nvN;
  bad |= fread(u.data(), sizeof(double), nvN, f) != nvN;
  bad |= fread(old.data(), sizeof(double), 2 * kN, f) != 2 * kN;
  bad |= fread(&g3old, sizeof(double), 1, f) != 1;
  if (bad) {
    fclose(f);
    return fail(ECHO_E_ARG, "echo_restart: %s ends early for this grid", path);
  }
  double th = 0.;
  int e = g->fo_on ? fo_dump_io(g, f, false, &th) : 0;
  fclose(f);
  if (e) return e;
  DevScalars S;
  HIPCHECK(hipMemcpyAsync(&S, g->S, sizeof S, hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  // out_restart, then system_metric and tout = t + dtout (echo.f90:164-172)
  S.t = t; S.timeold = timeold; S.timeinterval = timeinterval; S.g_cov_3_old = g3old;
  S.tout = t + S.dtout;
  S.run_crashed = 0; S.errcode = 0; S.temp_check = 0;
  S.crash_kind = CRASH_NONE; S.crash_cell = 0x7fffffff;
  HIPCHECK(hipMemcpyAsync(g->S, &S, sizeof S, hipMemcpyHostToDevice, g->stream));
  HIPCHECK(hipMemcpyAsync(g->V, v.data(), sizeof(double) * nvN, hipMemcpyHostToDevice, g->stream));
  HIPCHECK(hipMemcpyAsync(g->U[0], u.data(), sizeof(double) * nvN, hipMemcpyHostToDevice, g->stream));
  if (g->cfg.viscosity) {
    // vnewstep of the next step: the restored velocities; vold: the dump's
    const int par = S.vs_parity;
    HIPCHECK(hipMemcpyAsync(g->VS[par], v.data() + N, sizeof(double) * 3 * N, hipMemcpyHostToDevice, g->stream));
    HIPCHECK(hipMemcpyAsync(g->VS[par ^ 1], old.data() + N, sizeof(double) * 3 * N, hipMemcpyHostToDevice,
                            g->stream));
  }
  if (g->dump_track)
    HIPCHECK(hipMemcpyAsync(g->dump_buf[(S.iter - 1) & 1], old.data(), sizeof(double) * 2 * kN,
                            hipMemcpyHostToDevice, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  *iout = io;
  if (thyp) *thyp = th;
  return 0;
}

int echo_get_stop(echo_grid* g, int32_t* stopped, double* max_value) {
  if (!g) return fail(ECHO_E_ARG, "null argument");
  HIPCHECK(hipSetDevice(g->dev));
  DevScalars S;
  HIPCHECK(hipMemcpyAsync(&S, g->S, sizeof S, hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  if (stopped) *stopped = S.temp_check;
  if (max_value) *max_value = S.check_max;
  return 0;
}

int echo_set_tile(echo_grid* g, int32_t sweep, int32_t rows) {
  if (!g) return fail(ECHO_E_ARG, "null argument");
  if (g->next_stage != 1) return fail(ECHO_E_ARG, "echo_set_tile inside a step");
  return set_tile(g, sweep, rows);
}

int echo_get_tile(const echo_grid* g, int32_t* rows) {
  if (!g || !rows) return fail(ECHO_E_ARG, "null argument");
  for (int d = 0; d < 3; d++) rows[d] = g->tile[d];
  return 0;
}

int echo_tune(echo_grid* g, int32_t reps, double* ms) {
  if (!g) return fail(ECHO_E_ARG, "null argument");
  if (g->next_stage != 1) return fail(ECHO_E_ARG, "echo_tune inside a step");
  if (reps < 1) return fail(ECHO_E_ARG, "reps must be >= 1");
  HIPCHECK(hipSetDevice(g->dev));
  if (ms) for (int i = 0; i < 9; i++) ms[i] = 0.0;
  return tune_grid(g, reps, ms);
}

int echo_set_timing(echo_grid* g, int32_t on) {
  if (!g) return fail(ECHO_E_ARG, "null argument");
  g->timing = on != 0;
  return 0;
}

int echo_get_timing(echo_grid* g, double* t) {
  if (!g || !t) return fail(ECHO_E_ARG, "null argument");
  for (int i = 0; i < 4; i++) t[i] = g->tms[i];
  return 0;
}

int echo_ensemble_create(const echo_config* cfgs, int32_t n_events, const double* eos_rows,
                         int32_t n_rows, int32_t device, echo_ensemble** out) {
  if (!cfgs || !out || n_events < 1) return fail(ECHO_E_ARG, "echo_ensemble_create: invalid argument");
  *out = nullptr;
  HIPCHECK(hipSetDevice(device));
  echo_ensemble* e = new echo_ensemble;
  e->dev = device;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    delete e;
    return fail(ECHO_E_HIP, "stream creation failed");
  }
  size_t szF[3] = {0, 0, 0}, szL[3] = {0, 0, 0};
  long long szP = 0, maxN = 0, szUpre = 0;
  for (int i = 0; i < n_events; i++) {
    echo_grid* g = nullptr;
    int rc = grid_create(&cfgs[i], eos_rows, n_rows, device, e->stream, true, &g);
    if (rc) {
      release_ensemble(e);
      std::string msg = g_err;
      return fail(rc, "event %d: %s", i, msg.c_str());
    }
    g->graph_dirty = &e->dirty;
    e->g.push_back(g);
    for (int d = 0; d < 3; d++) {
      if (!g->active[d]) continue;
      szF[d] = std::max(szF[d], (size_t)((long long)face_slots(g->cfg, g->nv) * g->nf[d]));
      if (g->cfg.viscosity) szL[d] = std::max(szL[d], (size_t)(4LL * g->nf[d]));
    }
    szP = std::max(szP, g->partial_cap);
    maxN = std::max(maxN, g->N);
    szUpre = std::max(szUpre, g->fix_upre_cap);
  }
  for (int d = 0; d < 3; d++) {
    if (szF[d] && hipMalloc(&e->F[d], sizeof(double) * szF[d]) != hipSuccess) {
      release_ensemble(e);
      return fail(ECHO_E_NOMEM, "face buffer allocation failed");
    }
    if (szL[d] && hipMalloc(&e->L[d], sizeof(double) * szL[d]) != hipSuccess) {
      release_ensemble(e);
      return fail(ECHO_E_NOMEM, "face buffer allocation failed");
    }
  }
  if (hipMalloc(&e->partial, sizeof(double) * szP) != hipSuccess) {
    release_ensemble(e);
    return fail(ECHO_E_NOMEM, "partial buffer allocation failed");
  }
  // the con2prim repair scratch, one set for all events (they run one
  // after the other on one stream)
  if (hipMalloc(&e->fix.Pold, sizeof(double) * (size_t)maxN) != hipSuccess ||
      (szUpre && hipMalloc(&e->fix.Upre, sizeof(double) * (size_t)szUpre) != hipSuccess) ||
      hipMalloc(&e->fix.failbits, sizeof(unsigned) * (size_t)fix_words(maxN)) != hipSuccess ||
      hipMalloc(&e->fix.list, sizeof(int) * (size_t)maxN) != hipSuccess ||
      hipMalloc(&e->fix.level, sizeof(int) * (size_t)maxN) != hipSuccess ||
      hipMemset(e->fix.failbits, 0, sizeof(unsigned) * (size_t)fix_words(maxN)) != hipSuccess) {
    release_ensemble(e);
    return fail(ECHO_E_NOMEM, "repair scratch allocation failed");
  }
  for (echo_grid* g : e->g) {
    for (int d = 0; d < 3; d++) {
      g->F[d] = g->active[d] ? e->F[d] : nullptr;
      g->L[d] = (g->active[d] && g->cfg.viscosity) ? e->L[d] : nullptr;
    }
    g->partial = e->partial;
    g->fix = e->fix;
  }
  for (auto& v : e->ev) hipEventCreate(&v);
  *out = e;
  return 0;
}

void echo_ensemble_destroy(echo_ensemble* e) { release_ensemble(e); }

int echo_ensemble_size(const echo_ensemble* e) { return e ? (int)e->g.size() : 0; }

echo_grid* echo_ensemble_grid(echo_ensemble* e, int32_t i) {
  if (!e || i < 0 || i >= (int)e->g.size()) {
    fail(ECHO_E_ARG, "event index out of range");
    return nullptr;
  }
  return e->g[i];
}

int echo_ensemble_tune(echo_ensemble* e, int32_t reps) {
  if (!e || reps < 1) return fail(ECHO_E_ARG, "invalid argument");
  HIPCHECK(hipSetDevice(e->dev));
  for (echo_grid* g : e->g)
    if (int rc = tune_grid(g, reps, nullptr)) return rc;
  e->dirty = true;
  return 0;
}

int echo_ensemble_steps(echo_ensemble* e, int32_t nsteps, int32_t* n_crashed) {
  if (!e || nsteps < 0) return fail(ECHO_E_ARG, "invalid argument");
  HIPCHECK(hipSetDevice(e->dev));
  for (echo_grid* g : e->g)
    if (g->next_stage != 1) return fail(ECHO_E_ARG, "an event is inside a step");
  if (e->dirty || !e->gexec) {
    if (e->gexec) hipGraphExecDestroy(e->gexec);
    if (e->graph) hipGraphDestroy(e->graph);
    e->gexec = nullptr;
    e->graph = nullptr;
    HIPCHECK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    for (echo_grid* g : e->g) enqueue_step(g);
    hipGraph_t graph;
    HIPCHECK(hipStreamEndCapture(e->stream, &graph));
    e->graph = graph;
    HIPCHECK(hipGraphInstantiate(&e->gexec, graph, nullptr, nullptr, 0));
    e->dirty = false;
  }
  HIPCHECK(hipEventRecord(e->ev[0], e->stream));
  for (int s = 0; s < nsteps; s++) HIPCHECK(hipGraphLaunch(e->gexec, e->stream));
  HIPCHECK(hipEventRecord(e->ev[1], e->stream));
  HIPCHECK(hipEventSynchronize(e->ev[1]));
  float ms = 0;
  HIPCHECK(hipEventElapsedTime(&ms, e->ev[0], e->ev[1]));
  e->last_ms = ms;
  int nc = 0;
  for (echo_grid* g : e->g) nc += read_crash(g) != 0;
  if (n_crashed) *n_crashed = nc;
  return 0;
}

int echo_ensemble_last_ms(const echo_ensemble* e, double* ms) {
  if (!e || !ms) return fail(ECHO_E_ARG, "null argument");
  *ms = e->last_ms;
  return 0;
}

// ---------------------------------------------------------------------------
// output frames (out_vars, out.f90:232-591), written asynchronously
// ---------------------------------------------------------------------------
static int join_writer(echo_grid* g) {
  if (g->writer) {
    g->writer->join();
    delete g->writer;
    g->writer = nullptr;
  }
  const int st = g->writer_status;
  g->writer_status = 0;
  return st;
}

int echo_write_frame(echo_grid* g, const int32_t* sel, int32_t precision, int32_t iout, const char* dir) {
  if (!g || !sel || !dir) return fail(ECHO_E_ARG, "null argument");
  if (precision != 4 && precision != 8) return fail(ECHO_E_ARG, "precision must be 4 or 8");
  if (iout < 0 || iout > 9999) return fail(ECHO_E_ARG, "iout must be 0..9999");
  if (sel[27]) return fail(ECHO_E_ARG, "the comoving density (rc) is not written");
  const echo_config& c = g->cfg;
  const bool visc = c.viscosity != 0, mhd = c.mhd != 0;
  const bool anypit = sel[9] || sel[10] || sel[11] || sel[12];
  if (visc && sel[18] && (!c.obtained_zz || !anypit))
    return fail(ECHO_E_ARG, "pi^{eta eta} is written with TRACE_IMP 'zz' together with a pi^{t mu} field");
  if (int e = join_writer(g)) return e;
  HIPCHECK(hipSetDevice(g->dev));
  // file order of out_vars: (kind, index): kind 0 = primitive component,
  // kind 1 = derived field
  std::vector<std::pair<int, int>> order;
  const int prim5[5] = {KRH, KVX, KVY, KVZ, KPR};
  for (int q = 0; q < 5; q++) if (sel[q]) order.push_back({0, prim5[q]});
  for (int q = 0; q < 3; q++) if (sel[5 + q]) order.push_back({1, q});
  if (visc) {
    if (sel[8]) order.push_back({0, KPIBU});
    for (int q = 0; q < 4; q++) if (sel[9 + q]) order.push_back({1, 3 + q});
    const int pis[5] = {KPIXY, KPIXZ, KPIYZ, KPIXX, KPIYY};
    for (int q = 0; q < 5; q++) if (sel[13 + q]) order.push_back({0, pis[q]});
    if (sel[18]) order.push_back({1, 7});
  }
  if (sel[19]) order.push_back({1, 8});
  if (mhd) {
    // field slots of the layout (system.f90:92-124): resistive systems hold
    // B, E and psi after the fluid variables, and their E is written as
    // evolved (out.f90:404-413); ideal MHD derives E = -v x B (414-429)
    const int nf = visc ? 12 : 5;
    const int kb = c.mhd == 2 ? nf : KBX, kg = c.mhd == 2 ? nf + 6 : KGLM;
    for (int q = 0; q < 3; q++) if (sel[20 + q]) order.push_back({0, kb + q});
    for (int q = 0; q < 3; q++)
      if (sel[23 + q]) order.push_back(c.mhd == 2 ? std::make_pair(0, nf + 3 + q) : std::make_pair(1, 9 + q));
    if (c.divclean && sel[26]) order.push_back({0, kg});
  }
  const long long N = g->N;
  const size_t need = 1 + order.size() * (size_t)N;
  if (need > g->frame_cap) {
    if (g->frame_dev) hipFree(g->frame_dev);
    if (g->frame_host) hipHostFree(g->frame_host);
    g->frame_dev = nullptr;
    g->frame_host = nullptr;
    g->frame_cap = 0;
    if (hipMalloc(&g->frame_dev, sizeof(double) * (need + 1)) != hipSuccess ||
        hipHostMalloc(&g->frame_host, sizeof(double) * (need + 1)) != hipSuccess)
      return fail(ECHO_E_NOMEM, "frame buffers");
    g->frame_cap = need;
  }
  if (!g->copy_stream) {
    HIPCHECK(hipStreamCreateWithFlags(&g->copy_stream, hipStreamNonBlocking));
    HIPCHECK(hipEventCreateWithFlags(&g->frame_ready, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&g->frame_copied, hipEventDisableTiming));
  }
  // snapshot on the grid stream: t, the primitive arrays, the derived fields;
  // the extra slot after the arrays holds the EoS status of the derived pass
  double* fd = g->frame_dev;
  HIPCHECK(hipMemcpyAsync(fd, &g->S->t, sizeof(double), hipMemcpyDeviceToDevice, g->stream));
  OutSlots o;
  bool any_derived = false;
  for (int j = 0; j < 12; j++) o.dst[j] = -1;
  for (size_t s = 0; s < order.size(); s++) {
    if (order[s].first == 0)
      HIPCHECK(hipMemcpyAsync(fd + 1 + s * N, g->V + (size_t)order[s].second * N, sizeof(double) * N,
                              hipMemcpyDeviceToDevice, g->stream));
    else {
      o.dst[order[s].second] = (int)s;
      any_derived = true;
    }
  }
  int* fstat = reinterpret_cast<int*>(fd + need);
  HIPCHECK(hipMemsetAsync(fstat, 0, sizeof(int), g->stream));
  if (any_derived) {
    dispatch_sys(g->cfg, [&](auto sys, auto zz) {
      k_out_derived<decltype(sys)::value, decltype(zz)::value>
          <<<nblocks(N, UPDATE_BLOCK), UPDATE_BLOCK, 0, g->stream>>>(g->ga, g->V, o, fd, fstat);
    });
    HIPCHECK(hipGetLastError());
  }
  HIPCHECK(hipEventRecord(g->frame_ready, g->stream));
  HIPCHECK(hipStreamWaitEvent(g->copy_stream, g->frame_ready, 0));
  HIPCHECK(hipMemcpyAsync(g->frame_host, fd, sizeof(double) * (need + 1), hipMemcpyDeviceToHost, g->copy_stream));
  HIPCHECK(hipEventRecord(g->frame_copied, g->copy_stream));
  char name[32];
  snprintf(name, sizeof name, "/out%04d.dat", iout);
  const std::string path = std::string(dir) + name;
  const std::string tpath = std::string(dir) + "/time.dat";
  const size_t nslot = order.size();
  g->writer = new std::thread([g, path, tpath, precision, iout, nslot, N, need]() {
    hipSetDevice(g->dev);
    if (hipEventSynchronize(g->frame_copied) != hipSuccess) { g->writer_status = ECHO_E_HIP; return; }
    const double* h = g->frame_host;
    const int st = *reinterpret_cast<const int*>(h + need);
    if (st) { g->writer_status = st; return; }
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) { g->writer_status = ECHO_E_ARG; return; }
    bool ok = true;
    if (precision == 8) {
      ok = fwrite(h, sizeof(double), 1 + nslot * (size_t)N, f) == 1 + nslot * (size_t)N;
    } else {
      std::vector<float> buf(1 << 16);
      const size_t tot = 1 + nslot * (size_t)N;
      for (size_t i0 = 0; i0 < tot && ok; i0 += buf.size()) {
        const size_t m = std::min(buf.size(), tot - i0);
        for (size_t i = 0; i < m; i++) buf[i] = (float)h[i0 + i];
        ok = fwrite(buf.data(), sizeof(float), m, f) == m;
      }
    }
    ok = (fclose(f) == 0) && ok;
    FILE* tf = fopen(tpath.c_str(), "a");
    if (tf) {
      fprintf(tf, "%5d%16.9f\n", iout, h[0]);
      fclose(tf);
    }
    g->writer_status = ok ? 0 : ECHO_E_ARG;
  });
  return 0;
}

int echo_wait_frames(echo_grid* g) {
  if (!g) return fail(ECHO_E_ARG, "null argument");
  const int st = join_writer(g);
  if (st) return fail(st, "frame writer failed with code %d", st);
  return 0;
}

}  // extern "C"
