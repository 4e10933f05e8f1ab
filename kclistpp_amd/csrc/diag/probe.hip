// probe.hip - measurement helpers for the roofline figures (not part of the
// library).
//
//   probe copy <MiB>   plain device copy of <MiB> MiB (uint4 loads/stores),
//                      the reference kernel for scaling the FETCH_SIZE and
//                      WRITE_SIZE counters to bytes
//   probe lds          LDS read bandwidth of the whole GPU (ds_read_b128 from
//                      every lane of 4-wave workgroups), the LDS peak
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#define CHECK(x)                                                              \
  do {                                                                        \
    hipError_t e = (x);                                                       \
    if (e != hipSuccess) {                                                    \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));                  \
      return 1;                                                               \
    }                                                                         \
  } while (0)

__global__ void copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

constexpr int kLdsBytes = 64 * 1024;
constexpr int kLdsIters = 4096;

__global__ __launch_bounds__(256) void lds_read_kernel(uint4 *__restrict__ sink) {
  __shared__ uint4 buf[kLdsBytes / 16];
  const int tid = threadIdx.x;
  for (int i = tid; i < kLdsBytes / 16; i += 256) buf[i] = make_uint4(i, i ^ 1, i ^ 2, i ^ 3);
  __syncthreads();
  uint4 acc = make_uint4(0, 0, 0, 0);
  int idx = tid;
#pragma unroll 8
  for (int it = 0; it < kLdsIters; ++it) {
    const uint4 v = buf[idx];
    acc.x ^= v.x;
    acc.y ^= v.y;
    acc.z ^= v.z;
    acc.w ^= v.w;
    idx = (idx + 256) & (kLdsBytes / 16 - 1);
  }
  if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x12345678u) sink[blockIdx.x] = acc;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: probe copy <MiB> | probe lds\n");
    return 2;
  }
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  if (strcmp(argv[1], "copy") == 0) {
    size_t mib = argc > 2 ? strtoull(argv[2], nullptr, 10) : 1024;
    size_t bytes = mib << 20, n = bytes / 16;
    uint4 *src = nullptr, *dst = nullptr;
    CHECK(hipMalloc(&src, bytes));
    CHECK(hipMalloc(&dst, bytes));
    CHECK(hipMemset(src, 1, bytes));
    CHECK(hipMemset(dst, 0, bytes));
    copy_kernel<<<8192, 256>>>(src, dst, n);  // warm
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(a));
    copy_kernel<<<8192, 256>>>(src, dst, n);
    CHECK(hipEventRecord(b));
    CHECK(hipEventSynchronize(b));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, a, b));
    printf("copy %zu MiB: %.3f ms, %.1f GB/s (read+write), bytes read %zu written %zu\n", mib, ms,
           2.0 * bytes / (ms * 1e6), bytes, bytes);
    CHECK(hipFree(src));
    CHECK(hipFree(dst));
    return 0;
  }
  if (strcmp(argv[1], "lds") == 0) {
    int dev = 0;
    hipDeviceProp_t p;
    CHECK(hipGetDeviceProperties(&p, dev));
    const int blocks = p.multiProcessorCount * 8;
    uint4 *sink = nullptr;
    CHECK(hipMalloc(&sink, blocks * sizeof(uint4)));
    lds_read_kernel<<<blocks, 256>>>(sink);
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(a));
    for (int r = 0; r < 10; ++r) lds_read_kernel<<<blocks, 256>>>(sink);
    CHECK(hipEventRecord(b));
    CHECK(hipEventSynchronize(b));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, a, b));
    const double bytes = 10.0 * blocks * 256.0 * kLdsIters * 16.0;
    printf("lds read: %d CUs, %d blocks x 256 lanes x %d x 16 B, %.3f ms, %.1f GB/s, "
           "%.1f B/clk/CU at %d MHz\n",
           p.multiProcessorCount, blocks, kLdsIters, ms, bytes / (ms * 1e6),
           bytes / (ms * 1e-3) / (p.multiProcessorCount * (double)p.clockRate * 1e3),
           p.clockRate / 1000);
    CHECK(hipFree(sink));
    return 0;
  }
  fprintf(stderr, "unknown mode %s\n", argv[1]);
  return 2;
}
