// hbm_copy.hip -- streaming copy of a known byte count.
//
// Prints the achieved copy bandwidth and gives the HBM counters
// (FETCH_SIZE / WRITE_SIZE) a kernel of exactly known traffic, so that the
// counters read for the DEM kernels can be scaled to bytes.
//   hbm_copy [MiB per buffer, default 1024] [repeats, default 20]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void k_stream_copy(const double4* __restrict__ src,
                                                     double4* __restrict__ dst, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

int main(int argc, char** argv) {
  const long mib = argc > 1 ? std::atol(argv[1]) : 1024;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
  const long bytes = mib << 20;
  const long n = bytes / (long)sizeof(double4);
  double4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    std::fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  hipMemset(a, 0, bytes);
  hipMemset(b, 0, bytes);
  const int blocks = 256 * 8 * 4; // 4 blocks per SIMD group on 256 CUs
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
  hipDeviceSynchronize();
  hipEventRecord(e0, 0);
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
  hipEventRecord(e1, 0);
  hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  const double per = ms / reps;
  std::printf("{\"bytes_read_per_launch\": %ld, \"bytes_written_per_launch\": %ld, "
              "\"ms_per_launch\": %.6f, \"copy_GBps\": %.1f}\n",
              bytes, bytes, per, 2.0 * bytes / (per * 1e-3) / 1e9);
  hipFree(a);
  hipFree(b);
  return 0;
}
