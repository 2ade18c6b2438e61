// hbm_copy.hip -- streaming kernels of a known byte count.
//
// Gives the HBM counters (FETCH_SIZE / WRITE_SIZE) a kernel of exactly known
// traffic, so that the counters read for the DEM kernels can be scaled to
// bytes, and measures the practical bandwidth ceiling of this GPU:
//   k_stream_copy      grid-stride double4 copy (the counter calibration)
//   k_copy_unrolled    each thread keeps U double4 loads in flight before
//                      storing them (lane-coalesced, U blocks apart)
//   k_copy_unrolled_nt the same with non-temporal stores
//   k_read_sum         read-only: U loads in flight, one double per thread out
// Each variant runs `repeats` times on buffers of `MiB` each, at 4 to 64
// workgroups of 256 threads per CU; the JSON line gives every variant's
// GB/s (copy: bytes read + written) and the best copy.
//   hbm_copy [MiB per buffer, default 1024] [repeats, default 20]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void k_stream_copy(const double4* __restrict__ src,
                                                     double4* __restrict__ dst, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

template <int U, bool NT>
__global__ __launch_bounds__(256) void k_copy_unrolled(const double4* __restrict__ src,
                                                       double4* __restrict__ dst, long n) {
  const long span = (long)blockDim.x * U;
  for (long base = (long)blockIdx.x * span; base < n; base += (long)gridDim.x * span) {
    double4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long i = base + (long)u * blockDim.x + threadIdx.x;
      if (i < n) v[u] = src[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long i = base + (long)u * blockDim.x + threadIdx.x;
      if (i < n) {
        if (NT) {
          __builtin_nontemporal_store(v[u].x, &dst[i].x);
          __builtin_nontemporal_store(v[u].y, &dst[i].y);
          __builtin_nontemporal_store(v[u].z, &dst[i].z);
          __builtin_nontemporal_store(v[u].w, &dst[i].w);
        } else {
          dst[i] = v[u];
        }
      }
    }
  }
}

template <int U>
__global__ __launch_bounds__(256) void k_read_sum(const double4* __restrict__ src, double* __restrict__ out,
                                                  long n) {
  const long span = (long)blockDim.x * U;
  double s = 0.;
  for (long base = (long)blockIdx.x * span; base < n; base += (long)gridDim.x * span) {
    double4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long i = base + (long)u * blockDim.x + threadIdx.x;
      v[u] = i < n ? src[i] : make_double4(0., 0., 0., 0.);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) s += (v[u].x + v[u].y) + (v[u].z + v[u].w);
  }
  out[(long)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <class Launch>
static double time_ms(Launch launch, int reps) {
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  launch();
  hipDeviceSynchronize();
  hipEventRecord(e0, 0);
  for (int r = 0; r < reps; ++r) launch();
  hipEventRecord(e1, 0);
  hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return ms / reps;
}

int main(int argc, char** argv) {
  const long mib = argc > 1 ? std::atol(argv[1]) : 1024;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
  const long bytes = mib << 20;
  const long n = bytes / (long)sizeof(double4);
  int dev = 0, cus = 256;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  double4 *a = nullptr, *b = nullptr;
  double* sums = nullptr;
  const int grid = cus * 64; // the largest grid of the sweep (partial sums of the read kernel)
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
      hipMalloc(&sums, (size_t)grid * 256 * sizeof(double)) != hipSuccess) {
    std::fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  hipMemset(a, 0, bytes);
  hipMemset(b, 0, bytes);
  const double gs = time_ms([&] {
    hipLaunchKernelGGL(k_stream_copy, dim3(cus * 32), dim3(256), 0, 0, a, b, n);
  }, reps);
  auto copy_gbps = [&](double ms) { return 2.0 * bytes / (ms * 1e-3) / 1e9; };
  // sweep: grid-stride copy and unrolled copies at 4 to 64 workgroups per CU
  double best = copy_gbps(gs);
  std::printf("{\"bytes_read_per_launch\": %ld, \"bytes_written_per_launch\": %ld, "
              "\"ms_per_launch\": %.6f, \"copy_GBps\": %.1f, \"compute_units\": %d, \"sweep\": [",
              bytes, bytes, gs, copy_gbps(gs), cus);
  const int per_cu[] = {4, 8, 16, 32, 64};
  bool first = true;
  for (int m : per_cu) {
    const int g = cus * m;
    const double t_s = time_ms([&] {
      hipLaunchKernelGGL(k_stream_copy, dim3(g), dim3(256), 0, 0, a, b, n);
    }, reps);
    const double t_2 = time_ms([&] {
      hipLaunchKernelGGL((k_copy_unrolled<2, false>), dim3(g), dim3(256), 0, 0, a, b, n);
    }, reps);
    const double t_4 = time_ms([&] {
      hipLaunchKernelGGL((k_copy_unrolled<4, false>), dim3(g), dim3(256), 0, 0, a, b, n);
    }, reps);
    const double t_4nt = time_ms([&] {
      hipLaunchKernelGGL((k_copy_unrolled<4, true>), dim3(g), dim3(256), 0, 0, a, b, n);
    }, reps);
    const double t_r = time_ms([&] {
      hipLaunchKernelGGL((k_read_sum<8>), dim3(std::min(g, grid)), dim3(256), 0, 0, a, sums, n);
    }, reps);
    best = std::max(best, std::max(std::max(copy_gbps(t_s), copy_gbps(t_2)),
                                   std::max(copy_gbps(t_4), copy_gbps(t_4nt))));
    std::printf("%s{\"wg_per_cu\": %d, \"stream\": %.1f, \"unrolled2\": %.1f, \"unrolled4\": %.1f, "
                "\"unrolled4_nt\": %.1f, \"read\": %.1f}",
                first ? "" : ", ", m, copy_gbps(t_s), copy_gbps(t_2), copy_gbps(t_4), copy_gbps(t_4nt),
                bytes / (t_r * 1e-3) / 1e9);
    first = false;
  }
  std::printf("], \"best_copy_GBps\": %.1f}\n", best);
  hipFree(a);
  hipFree(b);
  hipFree(sums);
  return 0;
}
