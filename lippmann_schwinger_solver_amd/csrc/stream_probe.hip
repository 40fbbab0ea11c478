// Streaming-read ceiling of the device: the bandwidth a plain read-and-sum
// kernel reaches over a buffer of complex doubles, with non-temporal and with
// plain loads, over a sweep of grid sizes. It gives the practical ceiling
// beside the leaf M2L pass (one streamed operator, 16 B per entry) and the
// device's published peak. One JSON line per case on stdout.
//
//   stream_probe [bytes of the smaller buffer, default 443880000]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

typedef double v2 __attribute__((ext_vector_type(2)));
constexpr int kThreads = 256;
constexpr int kUnroll = 8;

#define PROBE_OK(expr)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      std::fprintf(stderr, "%s: %s\n", #expr, hipGetErrorString(e_));                   \
      return 1;                                                                         \
    }                                                                                   \
  } while (0)

// grid-stride sum of a[0, n); each thread writes its sum so nothing is elided
template <bool NT>
__global__ __launch_bounds__(kThreads) void read_kernel(const v2* __restrict__ a, size_t n, v2* __restrict__ out) {
  v2 acc = {0, 0};
  const size_t stride = size_t(gridDim.x) * kThreads;
  size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x;
  for (; i + (kUnroll - 1) * stride < n; i += kUnroll * stride) {
    v2 t[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) t[u] = NT ? __builtin_nontemporal_load(a + i + u * stride) : a[i + u * stride];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) acc += t[u];
  }
  for (; i < n; i += stride) acc += a[i];
  out[size_t(blockIdx.x) * kThreads + threadIdx.x] = acc;
}

template <bool NT>
int time_read(const v2* a, size_t n, v2* out, int grid, double* us) {
  hipEvent_t e0, e1;
  PROBE_OK(hipEventCreate(&e0));
  PROBE_OK(hipEventCreate(&e1));
  for (int w = 0; w < 3; ++w) hipLaunchKernelGGL((read_kernel<NT>), dim3(grid), dim3(kThreads), 0, 0, a, n, out);
  PROBE_OK(hipGetLastError());
  const int reps = 20;
  PROBE_OK(hipEventRecord(e0));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL((read_kernel<NT>), dim3(grid), dim3(kThreads), 0, 0, a, n, out);
  PROBE_OK(hipEventRecord(e1));
  PROBE_OK(hipEventSynchronize(e1));
  float ms = 0;
  PROBE_OK(hipEventElapsedTime(&ms, e0, e1));
  *us = 1e3 * ms / reps;
  PROBE_OK(hipEventDestroy(e0));
  PROBE_OK(hipEventDestroy(e1));
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const size_t small = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 443880000ull;
  const size_t sizes[2] = {small / sizeof(v2), (size_t(2) << 30) / sizeof(v2)};
  const int grids[5] = {1024, 2048, 4096, 8192, 16384};
  const size_t n_max = sizes[0] > sizes[1] ? sizes[0] : sizes[1];
  v2 *a = nullptr, *out = nullptr;
  PROBE_OK(hipMalloc((void**)&a, n_max * sizeof(v2)));
  PROBE_OK(hipMalloc((void**)&out, size_t(grids[4]) * kThreads * sizeof(v2)));
  PROBE_OK(hipMemset(a, 0, n_max * sizeof(v2)));
  PROBE_OK(hipDeviceSynchronize());
  for (size_t n : sizes)
    for (int g : grids)
      for (int nt = 1; nt >= 0; --nt) {
        double us = 0;
        if ((nt ? time_read<true>(a, n, out, g, &us) : time_read<false>(a, n, out, g, &us)) != 0) return 1;
        std::printf("{\"bytes\": %zu, \"grid\": %d, \"loads\": \"%s\", \"us\": %.2f, \"GBps\": %.1f}\n",
                    n * sizeof(v2), g, nt ? "nontemporal" : "plain", us, n * sizeof(v2) / (us * 1e-6) / 1e9);
        std::fflush(stdout);
      }
  PROBE_OK(hipFree(a));
  PROBE_OK(hipFree(out));
  return 0;
}
