// Device-side state of one grid (echo_grid) and host helpers shared by the
// translation units of the C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/echo_amd.h"
#include "kernels.hpp"

namespace echo {

inline thread_local std::string g_err;

inline int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHECK(x)                                                                     \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) return fail(ECHO_E_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

}  // namespace echo

struct echo_grid {
  using Eos = echo::Eos;
  using DevScalars = echo::DevScalars;
  using GridArgs = echo::GridArgs;
  echo_config cfg;
  int dev = 0;
  hipStream_t stream = nullptr;
  int nv = 0;
  long long N = 0;
  long long nf[3] = {0, 0, 0};
  int nb[3] = {0, 0, 0};
  bool active[3] = {false, false, false};
  double* V = nullptr;
  double* U[2] = {nullptr, nullptr};
  double* F[3] = {nullptr, nullptr, nullptr};
  double* L[3] = {nullptr, nullptr, nullptr};
  double* VS[2] = {nullptr, nullptr};
  echo::ImexBufs imex{};        // IMEX-SSP3(4,3,3) accumulators (INT_ALGO 4)
  double* partial = nullptr;
  double* tab = nullptr;
  int* tabix = nullptr;
  int n_tab = 0;
  DevScalars* S = nullptr;
  echo::FixBufs fix{};          // con2prim repair scratch (shared inside an ensemble)
  long long fix_upre_cap = 0;   // doubles of fix.Upre (0: no viscous smoothing)
  int* status = nullptr;
  GridArgs ga;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool timing = false;
  hipEvent_t ev[8] = {};
  double tms[4] = {0, 0, 0, 0};
  int next_stage = 1;
  int tile[3] = {1, 1, 1};    // flux tile shape (TY) per sweep
  bool owns_scratch = true;   // F, L, partial (shared inside an ensemble)
  bool owns_stream = true;
  long long partial_cap = 0;  // doubles in partial
  bool* graph_dirty = nullptr; // set when the launch shape changes (ensemble graph)
  bool stop_check = false;      // check_temp at the start of each step
  // asynchronous output frames
  double* frame_dev = nullptr;
  size_t frame_cap = 0;         // doubles
  double* frame_host = nullptr; // pinned
  hipStream_t copy_stream = nullptr;
  hipEvent_t frame_ready = nullptr, frame_copied = nullptr;
  std::thread* writer = nullptr;
  int writer_status = 0;
  // restart dump: v and u components krh..kvold_end at the start of the
  // last two steps (buffer iter & 1 is written when step iter starts), so
  // that vold / uold of out_dump are at hand between steps
  double* dump_buf[2] = {nullptr, nullptr};
  bool dump_track = false;
  // EoS table as converted at set-up (T, e, p, c_s^2 columns), host copy
  std::vector<double> tab_host;
  // freeze-out hypersurface: the freeze-out field of the older slice (fo_f[0]),
  // of the newer one (fo_f[1]) and a scratch field (fo_f[2]), the primitives
  // of the older slice, the per-cell crossing masks (reference loop order),
  // per-block counts, the elements of the last slice
  bool fo_on = false;
  bool fo_searched = false;     // the last slice was compared (not only stored)
  int fo_type = 0;
  double fo_value = 0., fo_t1 = 0., fo_t2 = 0., fo_hdr[2] = {0., 0.};
  double* fo_f[3] = {nullptr, nullptr, nullptr};
  double* fo_vb = nullptr;
  unsigned char* fo_mask = nullptr;
  long long* fo_bsum = nullptr;
  int* fo_status = nullptr;
  double* fo_rec = nullptr;
  long long fo_cap = 0, fo_n = 0;
  // learned con2prim inside the step (echo_set_c2p_net): the network (owned
  // by the caller), its Newton iterations and gate tolerance, and the
  // fallback count and list ([0] count, [1..N] cells)
  echo_mlp* net = nullptr;
  int net_iters = 3;
  double net_tol = 1e-12;
  int* net_scratch = nullptr;
};

// the learned inversion of a stage's result U into g->V after k_update
// (defined with the network kernels, c2p_mlp.hip)
void enqueue_net_c2p(echo_grid* g, const double* U);

// forget the grid's captured step (rebuilt at the next echo_steps)
inline void drop_graph(echo_grid* g) {
  if (g->gexec) hipGraphExecDestroy(g->gexec);
  if (g->graph) hipGraphDestroy(g->graph);
  g->gexec = nullptr;
  g->graph = nullptr;
}


namespace echo {

inline unsigned nblocks(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
// block count padded to a multiple of the XCD count (see xcd_block)
inline unsigned nblocks_xcd(long long n, int bs) {
  unsigned nb = nblocks(n, bs);
  return (nb + NXCD - 1) / NXCD * NXCD;
}

// the system of echo_config: hydro, viscous hydro ('no'/'zz'), ideal MHD
// without / with GLM, resistive MHD without / with viscosity
template <typename Fn>
inline void dispatch_sys(const echo_config& c, Fn&& fn) {
  if (c.mhd == 2) {
    if (!c.viscosity) fn(std::integral_constant<int, SYS_RMHD>(), std::false_type());
    else if (c.obtained_zz) fn(std::integral_constant<int, SYS_VRMHD>(), std::true_type());
    else fn(std::integral_constant<int, SYS_VRMHD>(), std::false_type());
  } else if (c.mhd) {
    if (c.divclean) fn(std::integral_constant<int, SYS_MHDGLM>(), std::false_type());
    else fn(std::integral_constant<int, SYS_MHD>(), std::false_type());
  } else if (c.viscosity) {
    if (c.obtained_zz) fn(std::integral_constant<int, SYS_VISC>(), std::true_type());
    else fn(std::integral_constant<int, SYS_VISC>(), std::false_type());
  } else {
    fn(std::integral_constant<int, SYS_HD>(), std::false_type());
  }
}

// find_pressure_from_temperature (eos.f90:1221-1252) on a host copy of the
// EoS, with the lower bisection bound started at 0 (the reference leaves it
// uninitialised); the bisection stops at |T(p) - T| <= 1e-10 T
inline double pressure_from_temperature(const Eos& E, double temp_input) {
  const double tolerance = 1.e-10;
  double en_try = (E.eqstate == 3) ? E.te[0] : 0.;
  if (en_try == 0.) en_try = tolerance / 2.;
  double pr_try, d1, d2, en, temp_try, s;
  eos_pressure(E, 1., en_try, pr_try, d1, d2);
  get_derived_data(E, 1., pr_try, en, temp_try, s);
  while (temp_try <= temp_input) {
    get_derived_data(E, 1., pr_try, en, temp_try, s);
    pr_try = pr_try * 1.1;
  }
  double pr_low = 0., pr_high = pr_try;
  while (fabs(temp_try - temp_input) > temp_input * tolerance) {
    pr_try = (pr_low + pr_high) / 2.;
    get_derived_data(E, 1., pr_try, en, temp_try, s);
    if (temp_try >= temp_input) pr_high = pr_try; else pr_low = pr_try;
  }
  return pr_try;
}

// the grid's EoS with host pointers into tab_host (plain bisection brackets)
inline Eos host_eos(const echo_grid* g) {
  Eos E = g->ga.eos;
  E.ixe = nullptr;
  E.ixp = nullptr;
  if (E.eqstate == 3) {
    const size_t n = (size_t)g->n_tab;
    const double* t = g->tab_host.data();
    E.tt = t; E.te = t + n; E.tp = t + 2 * n; E.tc = t + 3 * n;
  }
  return E;
}

}  // namespace echo
