// C-ABI of the batched con2prim paths (declared in include/echo_amd.h): the
// plain solver on device-resident states, and the learned inversion (network
// guess on MFMA, FP64 Newton refinement, residual gate, plain-solver
// fallback) with its on-device trainer.
#include <math.h>
#include <string.h>

#include <atomic>
#include <mutex>

#include "grid.hpp"
#include "mlp.hpp"

using namespace echo;

struct echo_mlp {
  int dev = 0;
  float* P = nullptr;       // FP32 parameters
  float* m1 = nullptr;      // Adam moments
  float* m2 = nullptr;
  float* G = nullptr;       // gradient
  float* Wq = nullptr;      // half-precision weights as float (backward pass)
  float* bias = nullptr;    // accumulator biases
  _Float16* packed = nullptr; // MFMA A operands
  // training scratch for `cap` states per batch
  long long cap = 0;
  float *ybuf = nullptr, *xbuf = nullptr, *zbuf = nullptr, *hbuf = nullptr, *dzbuf = nullptr;
  double* loss = nullptr;
  long long loss_cap = 0;
  // fallback list of the inversion
  int* fail_count = nullptr;
  int* fail_list = nullptr;
  long long fail_cap = 0;
};

namespace {

constexpr int C2P_BLOCK = 256;

// Persistent grid of k_mlp<MODE, NVU>: as many blocks as the device holds at
// once (the instantiation's occupancy, taken once), fewer for small n.  A
// larger grid would leave a second, partial round of blocks that each carry
// a full share of the tiles (1024 blocks against 768 resident at three waves
// per SIMD: the last 256 start when the first 768 end).
template <int MODE, int NVU>
int mlp_grid_blocks(int dev, long long n) {
  static std::atomic<int> per_cu{0};
  int per = per_cu.load(std::memory_order_relaxed);
  if (per < 1) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(&mlp::k_mlp<MODE, NVU>),
                                                     mlp::BLOCK, 0) != hipSuccess || per < 1)
      per = 1;
    per_cu.store(per, std::memory_order_relaxed);
  }
  int ncu = 256;
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const long long tiles = (n + 63) / 64;
  long long nb = (tiles + 3) / 4;
  const long long cap = (long long)per * ncu;
  if (nb > cap) nb = cap;
  if (nb < 1) nb = 1;
  return (int)nb;
}

void release_mlp(echo_mlp* m) {
  if (!m) return;
  hipSetDevice(m->dev);
  void* ptrs[] = {m->P, m->m1, m->m2, m->G, m->Wq, m->bias, m->packed, m->ybuf, m->xbuf,
                  m->zbuf, m->hbuf, m->dzbuf, m->loss, m->fail_count, m->fail_list};
  for (void* p : ptrs)
    if (p) hipFree(p);
  delete m;
}

int ensure_train_cap(echo_mlp* m, long long n) {
  if (n <= m->cap) return 0;
  void* ptrs[] = {m->ybuf, m->xbuf, m->zbuf, m->hbuf, m->dzbuf};
  for (void* p : ptrs)
    if (p) hipFree(p);
  m->ybuf = m->xbuf = m->zbuf = m->hbuf = m->dzbuf = nullptr;
  m->cap = 0;
  HIPCHECK(hipMalloc(&m->ybuf, sizeof(float) * mlp::NO * n));
  HIPCHECK(hipMalloc(&m->xbuf, sizeof(float) * mlp::NI * n));
  HIPCHECK(hipMalloc(&m->zbuf, sizeof(float) * 3 * mlp::NH * n));
  HIPCHECK(hipMalloc(&m->hbuf, sizeof(float) * 3 * mlp::NH * n));
  HIPCHECK(hipMalloc(&m->dzbuf, sizeof(float) * 4 * mlp::NH * n));
  m->cap = n;
  return 0;
}

int ensure_loss_cap(echo_mlp* m, long long n) {
  if (n <= m->loss_cap) return 0;
  if (m->loss) hipFree(m->loss);
  m->loss = nullptr;
  m->loss_cap = 0;
  HIPCHECK(hipMalloc(&m->loss, sizeof(double) * n));
  m->loss_cap = n;
  return 0;
}

int ensure_fail_cap(echo_mlp* m, long long n) {
  if (!m->fail_count) HIPCHECK(hipMalloc(&m->fail_count, sizeof(int)));
  if (n <= m->fail_cap) return 0;
  if (m->fail_list) hipFree(m->fail_list);
  m->fail_list = nullptr;
  m->fail_cap = 0;
  HIPCHECK(hipMalloc(&m->fail_list, sizeof(int) * n));
  m->fail_cap = n;
  return 0;
}

void enqueue_pack(echo_mlp* m, hipStream_t st) {
  mlp::k_mlp_pack<<<(mlp::PK_N + 255) / 256, 256, 0, st>>>(m->P, m->packed, m->bias, m->Wq);
}

mlp::NetArgs net_args(const echo_mlp* m) { return mlp::NetArgs{m->packed, m->bias}; }

// forward with saved activations, backward, weight gradient on states
// [off, off + n) of a data set of `stride` states
void enqueue_grad(echo_mlp* m, hipStream_t st, const double* du, const double* dv, long long stride,
                  long long off, long long n, double* loss_slot) {
  mlp::C2pArgs ca{};
  const int nb = mlp_grid_blocks<mlp::MODE_TRAIN, 8>(m->dev, n);
  mlp::k_mlp<mlp::MODE_TRAIN, 8><<<nb, mlp::BLOCK, 0, st>>>(net_args(m), ca, du, stride, off, n,
                                                            nullptr, nullptr, m->ybuf, m->xbuf,
                                                            m->zbuf, m->hbuf);
  mlp::k_mlp_bwd<<<nblocks(n, mlp::BWD_BLOCK), mlp::BWD_BLOCK, 0, st>>>(
      m->Wq, du, dv, stride, off, n, m->ybuf, m->zbuf, m->dzbuf, loss_slot);
  mlp::k_mlp_wgrad<<<64 * 3 + mlp::NO, 256, 0, st>>>(m->xbuf, m->hbuf, m->dzbuf, n, m->G);
}

bool surrogate_config_ok(const echo_config& c) {
  return c.mhd == 1 && c.algo_solver == 5 && c.eqstate == 1;
}

}  // namespace

// The learned inversion inside a grid's step (ideal MHD, after k_update
// wrote the stage's u and left V at the stage start): the network's guess,
// Newton and the gate for every cell, the plain solver from V for the cells
// the gate rejects, failures marked for the repair that follows.  The
// counter is cleared by a kernel, so the launches can be captured.
void enqueue_net_c2p(echo_grid* g, const double* U) {
  echo_mlp* m = g->net;
  int* cnt = g->net_scratch;
  const long long N = g->N;
  mlp::k_zero_int<<<1, 64, 0, g->stream>>>(cnt, 1);
  mlp::C2pArgs ca{g->cfg.coordinates, g->net_iters, g->net_tol, g->S, cnt, cnt + 1, 1};
  const int nb = g->cfg.divclean ? mlp_grid_blocks<mlp::MODE_C2P, 9>(g->dev, N)
                                  : mlp_grid_blocks<mlp::MODE_C2P, 8>(g->dev, N);
  int ncu = 256;
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, g->dev);
  if (g->cfg.divclean) {
    mlp::k_mlp<mlp::MODE_C2P, 9><<<nb, mlp::BLOCK, 0, g->stream>>>(net_args(m), ca, U, N, 0, N, g->V, nullptr,
                                                                   nullptr, nullptr, nullptr, nullptr);
    mlp::k_c2p_fallback_step<SYS_MHDGLM><<<ncu, 256, 0, g->stream>>>(g->ga, U, g->V, g->fix, cnt + 1, cnt);
  } else {
    mlp::k_mlp<mlp::MODE_C2P, 8><<<nb, mlp::BLOCK, 0, g->stream>>>(net_args(m), ca, U, N, 0, N, g->V, nullptr,
                                                                   nullptr, nullptr, nullptr, nullptr);
    mlp::k_c2p_fallback_step<SYS_MHD><<<ncu, 256, 0, g->stream>>>(g->ga, U, g->V, g->fix, cnt + 1, cnt);
  }
}

extern "C" {

int echo_mlp_create(int32_t device, echo_mlp** out) {
  if (!out) return fail(ECHO_E_ARG, "null argument");
  *out = nullptr;
  int ndev = 0;
  HIPCHECK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(ECHO_E_ARG, "device %d out of range", device);
  echo_mlp* m = new echo_mlp;
  m->dev = device;
  HIPCHECK(hipSetDevice(device));
  const size_t pb = sizeof(float) * mlp::NPARAM;
#define MALLOC(p, bytes)                                       \
  if (hipMalloc(&(p), (bytes)) != hipSuccess) {                \
    release_mlp(m);                                            \
    return fail(ECHO_E_NOMEM, "device allocation failed");     \
  }
  MALLOC(m->P, pb);
  MALLOC(m->m1, pb);
  MALLOC(m->m2, pb);
  MALLOC(m->G, pb);
  MALLOC(m->Wq, pb);
  MALLOC(m->bias, sizeof(float) * mlp::BS_N);
  MALLOC(m->packed, sizeof(_Float16) * mlp::PK_N);
#undef MALLOC
  HIPCHECK(hipMemset(m->P, 0, pb));
  HIPCHECK(hipMemset(m->m1, 0, pb));
  HIPCHECK(hipMemset(m->m2, 0, pb));
  HIPCHECK(hipMemset(m->G, 0, pb));
  enqueue_pack(m, nullptr);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipDeviceSynchronize());
  *out = m;
  return 0;
}

void echo_mlp_destroy(echo_mlp* m) { release_mlp(m); }

int echo_mlp_set_params(echo_mlp* m, const float* p, int64_t n) {
  if (!m || !p) return fail(ECHO_E_ARG, "null argument");
  if (n != mlp::NPARAM) return fail(ECHO_E_ARG, "expected %d parameters, got %lld", mlp::NPARAM, (long long)n);
  HIPCHECK(hipSetDevice(m->dev));
  const size_t pb = sizeof(float) * mlp::NPARAM;
  HIPCHECK(hipMemcpy(m->P, p, pb, hipMemcpyHostToDevice));
  HIPCHECK(hipMemset(m->m1, 0, pb));
  HIPCHECK(hipMemset(m->m2, 0, pb));
  enqueue_pack(m, nullptr);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipDeviceSynchronize());
  return 0;
}

int echo_mlp_get_params(echo_mlp* m, float* p, int64_t n) {
  if (!m || !p) return fail(ECHO_E_ARG, "null argument");
  if (n != mlp::NPARAM) return fail(ECHO_E_ARG, "expected %d parameters, got %lld", mlp::NPARAM, (long long)n);
  HIPCHECK(hipSetDevice(m->dev));
  HIPCHECK(hipMemcpy(p, m->P, sizeof(float) * mlp::NPARAM, hipMemcpyDeviceToHost));
  return 0;
}

int echo_mlp_forward(echo_mlp* m, echo_grid* g, const double* d_u, int64_t n, float* d_y) {
  if (!m || !g || !d_u || !d_y || n < 0) return fail(ECHO_E_ARG, "invalid argument");
  if (g->dev != m->dev) return fail(ECHO_E_ARG, "network and grid on different devices");
  if (n == 0) return 0;
  HIPCHECK(hipSetDevice(m->dev));
  mlp::C2pArgs ca{};
  mlp::k_mlp<mlp::MODE_FWD, 8><<<mlp_grid_blocks<mlp::MODE_FWD, 8>(m->dev, n), mlp::BLOCK, 0, g->stream>>>(
      net_args(m), ca, d_u, n, 0, n, nullptr, nullptr, d_y, nullptr, nullptr, nullptr);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(g->stream));
  return 0;
}

int echo_mlp_grad(echo_mlp* m, echo_grid* g, const double* d_u, const double* d_v, int64_t n,
                  float* grad, double* loss) {
  if (!m || !g || !d_u || !d_v || !grad || !loss || n <= 0) return fail(ECHO_E_ARG, "invalid argument");
  if (g->dev != m->dev) return fail(ECHO_E_ARG, "network and grid on different devices");
  HIPCHECK(hipSetDevice(m->dev));
  if (int e = ensure_train_cap(m, n)) return e;
  if (int e = ensure_loss_cap(m, 1)) return e;
  HIPCHECK(hipMemsetAsync(m->loss, 0, sizeof(double), g->stream));
  enqueue_grad(m, g->stream, d_u, d_v, n, 0, n, m->loss);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(grad, m->G, sizeof(float) * mlp::NPARAM, hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipMemcpyAsync(loss, m->loss, sizeof(double), hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  return 0;
}

int echo_mlp_train(echo_mlp* m, echo_grid* g, const double* d_u, const double* d_v, int64_t n,
                   const echo_mlp_train_cfg* cfg, float* loss_log) {
  if (!m || !g || !d_u || !d_v || !cfg) return fail(ECHO_E_ARG, "null argument");
  if (g->dev != m->dev) return fail(ECHO_E_ARG, "network and grid on different devices");
  if (cfg->steps <= 0 || cfg->batch <= 0 || cfg->batch > n)
    return fail(ECHO_E_ARG, "need steps > 0 and 0 < batch <= n");
  if (!(cfg->lr > 0.0f) || !(cfg->beta1 >= 0.0f && cfg->beta1 < 1.0f) ||
      !(cfg->beta2 >= 0.0f && cfg->beta2 < 1.0f) || !(cfg->eps > 0.0f))
    return fail(ECHO_E_ARG, "invalid optimiser settings");
  HIPCHECK(hipSetDevice(m->dev));
  const long long B = cfg->batch;
  if (int e = ensure_train_cap(m, B)) return e;
  if (int e = ensure_loss_cap(m, cfg->steps)) return e;
  HIPCHECK(hipMemsetAsync(m->loss, 0, sizeof(double) * cfg->steps, g->stream));
  const long long nbatch = n / B;
  const double pi = 3.14159265358979323846;
  for (int k = 0; k < cfg->steps; k++) {
    const long long off = (k % nbatch) * B;
    enqueue_grad(m, g->stream, d_u, d_v, n, off, B, m->loss + k);
    // cosine learning-rate schedule and Adam bias correction
    const double f = 0.5 * (1.0 + cos(pi * (double)k / (double)cfg->steps));
    const double lr = cfg->lr_final + (cfg->lr - cfg->lr_final) * f;
    const double t = (double)(k + 1);
    const double lr_t = lr * sqrt(1.0 - pow(cfg->beta2, t)) / (1.0 - pow(cfg->beta1, t));
    mlp::k_adam<<<(mlp::NPARAM + 255) / 256, 256, 0, g->stream>>>(m->P, m->m1, m->m2, m->G, (float)lr_t,
                                                                  cfg->beta1, cfg->beta2, cfg->eps);
    enqueue_pack(m, g->stream);
  }
  HIPCHECK(hipGetLastError());
  if (loss_log) {
    double* h = new double[cfg->steps];
    hipError_t e = hipMemcpyAsync(h, m->loss, sizeof(double) * cfg->steps, hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    for (int k = 0; k < cfg->steps && e == hipSuccess; k++) loss_log[k] = (float)h[k];
    delete[] h;
    if (e != hipSuccess) return fail(ECHO_E_HIP, "loss copy: %s", hipGetErrorString(e));
  }
  HIPCHECK(hipStreamSynchronize(g->stream));
  return 0;
}

int echo_cons2prim_batch_dev(echo_grid* g, const double* d_u, double* d_v, int32_t* d_status, int64_t n) {
  if (!g || !d_u || !d_v || !d_status || n < 0) return fail(ECHO_E_ARG, "invalid argument");
  if (n == 0) return 0;
  HIPCHECK(hipSetDevice(g->dev));
  if (g->timing) HIPCHECK(hipEventRecord(g->ev[0], g->stream));
  dispatch_sys(g->cfg, [&](auto sys, auto zz) {
    k_c2p_batch<decltype(sys)::value, decltype(zz)::value>
        <<<nblocks(n, UPDATE_BLOCK), UPDATE_BLOCK, 0, g->stream>>>(g->ga, d_u, d_v, d_status, n);
  });
  HIPCHECK(hipGetLastError());
  if (g->timing) HIPCHECK(hipEventRecord(g->ev[1], g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  if (g->timing) {
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, g->ev[0], g->ev[1]));
    g->tms[0] = ms;
    g->tms[1] = g->tms[2] = g->tms[3] = 0.0;
  }
  return 0;
}

// the learned inversion with the fallback counter and list at fail_count,
// fail_list (room for n entries); shared by both entry points below
static int surrogate_run(echo_grid* g, echo_mlp* m, const double* d_u, double* d_v, int32_t* d_status,
                         int64_t n, int32_t iters, double tol, int64_t* n_fallback, int* fail_count,
                         int* fail_list) {
  mlp::k_zero_int<<<1, 64, 0, g->stream>>>(fail_count, 1);
  mlp::C2pArgs ca{g->cfg.coordinates, iters, tol, g->S, fail_count, fail_list, 0};
  const int nb = g->cfg.divclean ? mlp_grid_blocks<mlp::MODE_C2P, 9>(m->dev, n)
                                  : mlp_grid_blocks<mlp::MODE_C2P, 8>(m->dev, n);
  if (g->timing) HIPCHECK(hipEventRecord(g->ev[0], g->stream));
  if (g->cfg.divclean)
    mlp::k_mlp<mlp::MODE_C2P, 9><<<nb, mlp::BLOCK, 0, g->stream>>>(net_args(m), ca, d_u, n, 0, n, d_v,
                                                                   d_status, nullptr, nullptr, nullptr, nullptr);
  else
    mlp::k_mlp<mlp::MODE_C2P, 8><<<nb, mlp::BLOCK, 0, g->stream>>>(net_args(m), ca, d_u, n, 0, n, d_v,
                                                                   d_status, nullptr, nullptr, nullptr, nullptr);
  HIPCHECK(hipGetLastError());
  if (g->timing) HIPCHECK(hipEventRecord(g->ev[1], g->stream));
  int ncu = 256;
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, g->dev);
  if (g->cfg.divclean)
    mlp::k_c2p_fallback<SYS_MHDGLM><<<ncu, 256, 0, g->stream>>>(g->ga, d_u, d_v, d_status, n, fail_list,
                                                                fail_count);
  else
    mlp::k_c2p_fallback<SYS_MHD><<<ncu, 256, 0, g->stream>>>(g->ga, d_u, d_v, d_status, n, fail_list,
                                                             fail_count);
  HIPCHECK(hipGetLastError());
  if (g->timing) HIPCHECK(hipEventRecord(g->ev[2], g->stream));
  int nf = 0;
  HIPCHECK(hipMemcpyAsync(&nf, fail_count, sizeof(int), hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  if (n_fallback) *n_fallback = nf;
  if (g->timing) {
    float a = 0, b = 0;
    HIPCHECK(hipEventElapsedTime(&a, g->ev[0], g->ev[1]));
    HIPCHECK(hipEventElapsedTime(&b, g->ev[1], g->ev[2]));
    g->tms[0] = a;
    g->tms[1] = b;
    g->tms[2] = g->tms[3] = 0.0;
  }
  return 0;
}

static int surrogate_check(echo_grid* g, echo_mlp* m, const double* d_u, double* d_v, int32_t* d_status,
                           int64_t n, int32_t iters, double tol) {
  if (!g || !m || !d_u || !d_v || !d_status || n < 0) return fail(ECHO_E_ARG, "invalid argument");
  if (g->dev != m->dev) return fail(ECHO_E_ARG, "network and grid on different devices");
  if (!surrogate_config_ok(g->cfg))
    return fail(ECHO_E_ARG, "the learned inversion covers ideal MHD with p = e/3 (mhd 1, algo_solver 5, eqstate 1)");
  if (iters < 0 || iters > 8 || !(tol > 0.0)) return fail(ECHO_E_ARG, "need 0 <= iters <= 8 and tol > 0");
  if (n > 0x7fffffffLL) return fail(ECHO_E_ARG, "at most 2^31-1 states per call");
  return 0;
}

int echo_set_c2p_net(echo_grid* g, echo_mlp* m, int32_t iters, double tol) {
  if (!g) return fail(ECHO_E_ARG, "null argument");
  HIPCHECK(hipSetDevice(g->dev));
  if (g->next_stage != 1) return fail(ECHO_E_ARG, "echo_set_c2p_net: the grid is inside a step");
  if (!m) {
    g->net = nullptr;
    g->ga.ext_c2p = 0;
    hipFree(g->net_scratch);
    g->net_scratch = nullptr;
  } else {
    if (g->dev != m->dev) return fail(ECHO_E_ARG, "network and grid on different devices");
    if (!surrogate_config_ok(g->cfg))
      return fail(ECHO_E_ARG, "the learned inversion covers ideal MHD with p = e/3 (mhd 1, algo_solver 5, eqstate 1)");
    if (g->cfg.nrk == 4)
      return fail(ECHO_E_ARG, "the learned inversion runs in the RK2 / RK3 steps (not IMEX-SSP3(4,3,3))");
    if (iters < 0 || iters > 8 || !(tol > 0.0)) return fail(ECHO_E_ARG, "need 0 <= iters <= 8 and tol > 0");
    if (g->N > 0x7fffffffLL) return fail(ECHO_E_ARG, "at most 2^31-1 cells");
    if (!g->net_scratch && hipMalloc(&g->net_scratch, sizeof(int) * (size_t)(g->N + 1)) != hipSuccess) {
      g->net_scratch = nullptr;
      return fail(ECHO_E_NOMEM, "echo_set_c2p_net: device allocation failed");
    }
    g->net = m;
    g->net_iters = iters;
    g->net_tol = tol;
    g->ga.ext_c2p = 1;
    // take the network kernel's occupancy now, outside any graph capture
    if (g->cfg.divclean) mlp_grid_blocks<mlp::MODE_C2P, 9>(g->dev, 1);
    else mlp_grid_blocks<mlp::MODE_C2P, 8>(g->dev, 1);
  }
  drop_graph(g);
  if (g->graph_dirty) *g->graph_dirty = true;
  return 0;
}

int echo_get_c2p_fallbacks(echo_grid* g, int64_t* n) {
  if (!g || !n) return fail(ECHO_E_ARG, "null argument");
  HIPCHECK(hipSetDevice(g->dev));
  echo::DevScalars S;
  HIPCHECK(hipMemcpyAsync(&S, g->S, sizeof S, hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  *n = S.nfallback;
  return 0;
}

int echo_c2p_surrogate(echo_grid* g, echo_mlp* m, const double* d_u, double* d_v, int32_t* d_status,
                       int64_t n, int32_t iters, double tol, int64_t* n_fallback) {
  if (int e = surrogate_check(g, m, d_u, d_v, d_status, n, iters, tol)) return e;
  if (n_fallback) *n_fallback = 0;
  if (n == 0) return 0;
  HIPCHECK(hipSetDevice(g->dev));
  if (int e = ensure_fail_cap(m, n)) return e;
  return surrogate_run(g, m, d_u, d_v, d_status, n, iters, tol, n_fallback, m->fail_count, m->fail_list);
}

}  // extern "C"

// ---- the allocation-free variant of echo_c2p_surrogate --------------------
// Its scratch (the fallback counter in word 0, the fallback list after it)
// is this file-scope device array, one copy per device, held until the
// process ends.  On the host the name is a stand-in, not a device address:
// scratch_on() takes the address of each device's copy once, with
// hipGetSymbolAddress while that device is current, and every use goes
// through that address.
static __device__ __align__(256) int scratch[1 + ECHO_NOALLOC_MAX_STATES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
int* scratch_dev[MAX_DEVICES] = {};

int scratch_on(int dev, int** out) {
  if (dev < 0 || dev >= MAX_DEVICES) return fail(ECHO_E_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(ECHO_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    scratch_dev[dev] = static_cast<int*>(p);
  }
  *out = scratch_dev[dev];
  return 0;
}
}  // namespace

extern "C" {

int echo_c2p_surrogate_noalloc(echo_grid* g, echo_mlp* m, const double* d_u, double* d_v, int32_t* d_status,
                               int64_t n, int32_t iters, double tol, int64_t* n_fallback) {
  if (int e = surrogate_check(g, m, d_u, d_v, d_status, n, iters, tol)) return e;
  if (n > ECHO_NOALLOC_MAX_STATES)
    return fail(ECHO_E_ARG, "echo_c2p_surrogate_noalloc: at most %d states per call (ECHO_NOALLOC_MAX_STATES), "
                            "got %lld", ECHO_NOALLOC_MAX_STATES, (long long)n);
  if (n_fallback) *n_fallback = 0;
  if (n == 0) return 0;
  HIPCHECK(hipSetDevice(g->dev));
  int* sc = nullptr;
  if (int e = scratch_on(g->dev, &sc)) return e;
  return surrogate_run(g, m, d_u, d_v, d_status, n, iters, tol, n_fallback, sc, sc + 1);
}

// ---- device memory for callers without a runtime of their own ------------
int echo_dev_alloc(int32_t device, int64_t bytes, void** ptr) {
  if (!ptr || bytes < 0) return fail(ECHO_E_ARG, "invalid argument");
  *ptr = nullptr;
  HIPCHECK(hipSetDevice(device));
  if (bytes == 0) return 0;
  if (hipMalloc(ptr, (size_t)bytes) != hipSuccess) {
    *ptr = nullptr;
    return fail(ECHO_E_NOMEM, "hipMalloc of %lld bytes failed", (long long)bytes);
  }
  return 0;
}

int echo_dev_free(int32_t device, void* ptr) {
  HIPCHECK(hipSetDevice(device));
  if (ptr) HIPCHECK(hipFree(ptr));
  return 0;
}

int echo_dev_upload(int32_t device, void* dst, const void* src, int64_t bytes) {
  if ((!dst || !src) && bytes > 0) return fail(ECHO_E_ARG, "null argument");
  HIPCHECK(hipSetDevice(device));
  if (bytes > 0) HIPCHECK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
  return 0;
}

int echo_dev_download(int32_t device, void* dst, const void* src, int64_t bytes) {
  if ((!dst || !src) && bytes > 0) return fail(ECHO_E_ARG, "null argument");
  HIPCHECK(hipSetDevice(device));
  if (bytes > 0) HIPCHECK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
