// Device-to-device copies of a known byte count.
//
//   copy_calib [MiB]              plain grid-stride copy, launched 3 times.
//       Profiler counters (FETCH_SIZE, WRITE_SIZE) read for this kernel give
//       the factor between counter units and bytes moved, used to scale the
//       counters of the grid-update kernels.
//   copy_calib --bandwidth [MiB]  times copy kernels of several shapes with
//       HIP events (best of 5) and prints the read + write rate of each: the
//       HBM bandwidth a streaming kernel can reach on this device.
//
// MiB defaults to 1024 (bandwidth mode: 2048); it must be a multiple of 64.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef double d2 __attribute__((ext_vector_type(2)));

__global__ void k_copy(const d2* __restrict__ a, d2* __restrict__ b, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

// U 16-byte elements per thread per pass, optionally non-temporal; the
// caller makes n a multiple of U * 256 * gridDim.x
template <int U, bool NT>
__global__ __launch_bounds__(256) void k_copy_wide(const d2* __restrict__ a, d2* __restrict__ b, long long n) {
  const long long step = (long long)gridDim.x * 256 * U;
  for (long long base = (long long)blockIdx.x * 256 * U; base < n; base += step) {
    d2 r[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const d2* p = &a[base + u * 256 + threadIdx.x];
      r[u] = NT ? __builtin_nontemporal_load(p) : *p;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      d2* q = &b[base + u * 256 + threadIdx.x];
      if (NT) __builtin_nontemporal_store(r[u], q);
      else *q = r[u];
    }
  }
}

template <typename F>
static float best_ms(F launch) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1.f;
  launch();
  float best = 1e30f;
  for (int r = 0; r < 5; r++) {
    float ms = 0.f;
    if (hipEventRecord(e0) != hipSuccess) return -1.f;
    launch();
    if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      return -1.f;
    if (ms < best) best = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

int main(int argc, char** argv) {
  const bool bw = argc > 1 && strcmp(argv[1], "--bandwidth") == 0;
  const char* size_arg = bw ? (argc > 2 ? argv[2] : nullptr) : (argc > 1 ? argv[1] : nullptr);
  long long mib = size_arg ? atoll(size_arg) : (bw ? 2048 : 1024);
  if (mib <= 0 || mib % 64 != 0) {
    fprintf(stderr, "copy_calib: MiB must be a positive multiple of 64\n");
    return 1;
  }
  long long bytes = mib << 20;
  long long n = bytes / 16;
  d2 *a, *b;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    fprintf(stderr, "copy_calib: allocation failed\n");
    return 1;
  }
  if (hipMemset(a, 0, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) return 1;
  if (!bw) {
    for (int r = 0; r < 3; r++) k_copy<<<4096, 256>>>(a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    printf("copy_calib: %lld bytes read and %lld bytes written per launch, 3 launches\n", bytes, bytes);
  } else {
    // n = mib * 2^16 elements: a multiple of 4 * 256 * 64 for every grid below
    const double gb = 2.0 * (double)bytes / 1e9;
    struct { const char* name; float ms; } r[5];
    r[0] = {"grid-stride, 4096 x 256", best_ms([&] { k_copy<<<4096, 256>>>(a, b, n); })};
    r[1] = {"4 x 16 B per thread, 2048 blocks", best_ms([&] { k_copy_wide<4, false><<<2048, 256>>>(a, b, n); })};
    r[2] = {"4 x 16 B non-temporal, 2048 blocks", best_ms([&] { k_copy_wide<4, true><<<2048, 256>>>(a, b, n); })};
    r[3] = {"1 x 16 B non-temporal, n / 256 blocks",
            best_ms([&] { k_copy_wide<1, true><<<(unsigned)(n / 256), 256>>>(a, b, n); })};
    r[4] = {"hipMemcpyAsync device to device",
            best_ms([&] { (void)hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice); })};
    printf("copy of %lld MiB (read + write %.3f GB), best of 5\n", mib, gb);
    for (auto& x : r) {
      if (x.ms <= 0.f) { fprintf(stderr, "copy_calib: timing failed\n"); return 1; }
      printf("%-40s %8.3f ms %8.0f GB/s\n", x.name, x.ms, gb / x.ms * 1e3);
    }
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
