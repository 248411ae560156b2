// hbm_probe.hip -- achievable HBM bandwidth of the GPU for the access shapes
// the time-stepping kernels use: float4 read-only, write-only, copy (plain and
// non-temporal), 3 reads + 1 write, and a one-float-per-lane copy.  Each
// shape is timed with HIP events over a sweep of grid and block sizes; the
// best rate is printed, in GB/s (10^9 bytes).  4 GiB per buffer, 4 buffers.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                  \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(1);                                                            \
    }                                                                          \
  } while (0)

typedef float v4 __attribute__((ext_vector_type(4)));

__global__ void k_read(const v4 *__restrict__ a, size_t n4, float *__restrict__ out) {
  float s = 0.f;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4;
       i += (size_t)gridDim.x * blockDim.x) {
    const v4 v = a[i];
    s += v.x + v.y + v.z + v.w;
  }
  if (s == 12345.678f) out[0] = s;  // keeps the loads; never true for zeros
}
__global__ void k_write(v4 *__restrict__ a, size_t n4) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4;
       i += (size_t)gridDim.x * blockDim.x)
    a[i] = v4{1.f, 2.f, 3.f, 4.f};
}
__global__ void k_copy(const v4 *__restrict__ a, v4 *__restrict__ b, size_t n4) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4;
       i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}
__global__ void k_copy_nt(const v4 *__restrict__ a, v4 *__restrict__ b, size_t n4) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4;
       i += (size_t)gridDim.x * blockDim.x)
    __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}
__global__ void k_r3w1(const v4 *__restrict__ a, const v4 *__restrict__ b,
                       const v4 *__restrict__ c, v4 *__restrict__ d, size_t n4) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4;
       i += (size_t)gridDim.x * blockDim.x)
    d[i] = a[i] + b[i] + c[i];
}
__global__ void k_copy1(const float *__restrict__ a, float *__restrict__ b, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

int main() {
  const size_t bytes = (size_t)4 << 30;
  const size_t n4 = bytes / 16;
  v4 *a, *b, *c, *d;
  float *out;
  CK(hipMalloc(&a, bytes)); CK(hipMalloc(&b, bytes));
  CK(hipMalloc(&c, bytes)); CK(hipMalloc(&d, bytes));
  CK(hipMalloc(&out, 64));
  CK(hipMemset(a, 0, bytes)); CK(hipMemset(b, 0, bytes));
  CK(hipMemset(c, 0, bytes)); CK(hipMemset(d, 0, bytes));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
  const int grids[] = {1024, 2048, 4096, 8192, 16384, 65536};
  const int blocks[] = {256, 512, 1024};
  auto best = [&](const char *name, double moved, auto launch) {
    double top = 0;
    int bg = 0, bb = 0;
    for (int g : grids)
      for (int t : blocks) {
        launch(g, t);
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        for (int r = 0; r < 5; ++r) launch(g, t);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms;
        CK(hipEventElapsedTime(&ms, e0, e1));
        const double gbs = moved * bytes / (ms / 5 * 1e-3) / 1e9;
        if (gbs > top) { top = gbs; bg = g; bb = t; }
      }
    std::printf("%-24s %6.0f GB/s  (grid %d x %d)\n", name, top, bg, bb);
    std::fflush(stdout);
  };
  best("read float4", 1, [&](int g, int t) { hipLaunchKernelGGL(k_read, dim3(g), dim3(t), 0, 0, a, n4, out); });
  best("write float4", 1, [&](int g, int t) { hipLaunchKernelGGL(k_write, dim3(g), dim3(t), 0, 0, b, n4); });
  best("copy float4", 2, [&](int g, int t) { hipLaunchKernelGGL(k_copy, dim3(g), dim3(t), 0, 0, a, b, n4); });
  best("copy float4 nontemporal", 2, [&](int g, int t) { hipLaunchKernelGGL(k_copy_nt, dim3(g), dim3(t), 0, 0, a, b, n4); });
  best("3 reads + 1 write", 4, [&](int g, int t) { hipLaunchKernelGGL(k_r3w1, dim3(g), dim3(t), 0, 0, a, b, c, d, n4); });
  const size_t n = bytes / 4;
  for (int t : {256, 512}) {
    const unsigned g = (unsigned)((n + t - 1) / t);
    auto one = [&]() { hipLaunchKernelGGL(k_copy1, dim3(g), dim3(t), 0, 0, (const float *)a, (float *)b, n); };
    one();
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(e0));
    for (int r = 0; r < 5; ++r) one();
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms;
    CK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("copy 1 float/lane, %3d   %6.0f GB/s\n", t, 2.0 * bytes / (ms / 5 * 1e-3) / 1e9);
  }
  CK(hipFree(a)); CK(hipFree(b)); CK(hipFree(c)); CK(hipFree(d)); CK(hipFree(out));
  return 0;
}
