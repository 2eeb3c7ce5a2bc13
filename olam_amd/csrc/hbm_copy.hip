// hbm_copy.hip -- plain device-to-device copy of a known number of bytes.
//
// Used (1) to measure the achievable HBM copy bandwidth of the GPU the library runs
// on, reported next to the roofline fractions, and (2) as the calibration kernel for
// rocprofv3 FETCH_SIZE / WRITE_SIZE counters: the copy reads and writes exactly
// `bytes` each, so the counters' reading for it gives their scale.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void k_hbm_copy(const float4 *__restrict__ a, float4 *__restrict__ b, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

// one-shot variant: each thread moves 4 float4 with all loads issued before the stores
__global__ __launch_bounds__(256) void k_hbm_copy4(const float4 *__restrict__ a, float4 *__restrict__ b, long n) {
  long i = ((long)blockIdx.x * blockDim.x) * 4 + threadIdx.x;
  float4 v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) { long j = i + (long)u * blockDim.x; if (j < n) v[u] = a[j]; }
#pragma unroll
  for (int u = 0; u < 4; ++u) { long j = i + (long)u * blockDim.x; if (j < n) b[j] = v[u]; }
}

#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e)); return 1; } } while (0)

int main(int argc, char **argv) {
  long bytes = (argc > 1) ? std::atol(argv[1]) : (1L << 30);
  int reps = (argc > 2) ? std::atoi(argv[2]) : 10;
  long n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  CK(hipMalloc(&a, n * 16));
  CK(hipMalloc(&b, n * 16));
  CK(hipMemset(a, 0, n * 16));
  CK(hipMemset(b, 0, n * 16));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  int blocks = 256 * 8 * 4;     // 8 blocks of 256 threads per CU x 4, grid-stride
  hipLaunchKernelGGL(k_hbm_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
  CK(hipDeviceSynchronize());
  float best = 1e30f, best4 = 1e30f;
  unsigned blocks4 = (unsigned)((n + 1023) / 1024);
  hipLaunchKernelGGL(k_hbm_copy4, dim3(blocks4), dim3(256), 0, 0, a, b, n);
  CK(hipDeviceSynchronize());
  for (int r = 0; r < reps; ++r) {
    float ms = 0.f;
    CK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(k_hbm_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) best = ms;
    CK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(k_hbm_copy4, dim3(blocks4), dim3(256), 0, 0, a, b, n);
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best4) best4 = ms;
  }
  float b1 = best < best4 ? best : best4;
  std::printf("{\"copy_bytes\": %ld, \"best_ms\": %.4f, \"copy_GBps\": %.1f, \"gridstride_GBps\": %.1f, \"oneshot_GBps\": %.1f}\n",
              n * 16, b1, 2.0 * n * 16 / (b1 * 1e-3) / 1e9, 2.0 * n * 16 / (best * 1e-3) / 1e9,
              2.0 * n * 16 / (best4 * 1e-3) / 1e9);
  CK(hipFree(a));
  CK(hipFree(b));
  return 0;
}
