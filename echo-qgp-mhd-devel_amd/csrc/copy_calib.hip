// Plain device-to-device copy of a known byte count, launched 3 times.
// Profiler counters (FETCH_SIZE, WRITE_SIZE) read for this kernel give the
// factor between counter units and bytes moved, used to scale the counters
// of the grid-update kernels.  Usage: copy_calib [MiB]  (default 1024)
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

__global__ void k_copy(const double2* __restrict__ a, double2* __restrict__ b, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

int main(int argc, char** argv) {
  long long mib = argc > 1 ? atoll(argv[1]) : 1024;
  long long bytes = mib << 20;
  long long n = bytes / 16;
  double2 *a, *b;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    fprintf(stderr, "copy_calib: allocation failed\n");
    return 1;
  }
  if (hipMemset(a, 0, bytes) != hipSuccess) return 1;
  for (int r = 0; r < 3; r++) k_copy<<<4096, 256>>>(a, b, n);
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  printf("copy_calib: %lld bytes read and %lld bytes written per launch, 3 launches\n", bytes, bytes);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
