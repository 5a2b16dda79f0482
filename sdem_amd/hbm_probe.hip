// hbm_probe -- the HBM rate this GPU reaches with plain streaming kernels,
// the ceiling that the DEM kernels' measured traffic is compared with.
// Copies 1 GiB between two device buffers (16-byte and 8-byte elements,
// with and without nontemporal hints), reads it, writes it, and times the
// runtime's own device-to-device copy for comparison.  Rates count bytes
// read plus bytes written per second (HIP events, 20 launches after 3).
//
//   sdem_amd/hbm_probe          (built by __graft_entry__.build())
#include <hip/hip_runtime.h>

#include <cstdio>

#define CK(x)                                                                  \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      return 1;                                                                \
    }                                                                          \
  } while (0)

__global__ void k_copy16(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_copy16_nt(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* s = (const float*)(a + i);
  float* d = (float*)(b + i);
  const float x = __builtin_nontemporal_load(s), y = __builtin_nontemporal_load(s + 1);
  const float z = __builtin_nontemporal_load(s + 2), w = __builtin_nontemporal_load(s + 3);
  __builtin_nontemporal_store(x, d);
  __builtin_nontemporal_store(y, d + 1);
  __builtin_nontemporal_store(z, d + 2);
  __builtin_nontemporal_store(w, d + 3);
}

__global__ void k_copy8(const double* __restrict__ a, double* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_read16(const float4* __restrict__ a, float* __restrict__ out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  float s = 0.0f;
  if (i < n) {
    const float4 v = a[i];
    s = v.x + v.y + v.z + v.w;
  }
  if (s == 12345.678f) out[0] = s;  // never true for the zero-filled input; keeps the load
}

__global__ void k_write16(float4* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = make_float4(1.0f, 2.0f, 3.0f, 4.0f);
}

int main() {
  const size_t bytes = (size_t)1 << 30, n16 = bytes / 16, n8 = bytes / 8;
  const unsigned g16 = (unsigned)((n16 + 255) / 256), g8 = (unsigned)((n8 + 255) / 256);
  float4 *a = nullptr, *b = nullptr;
  float* o = nullptr;
  CK(hipMalloc(&a, bytes));
  CK(hipMalloc(&b, bytes));
  CK(hipMalloc(&o, 64));
  CK(hipMemset(a, 0, bytes));
  CK(hipMemset(b, 0, bytes));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  auto rate = [&](const char* name, double moved, auto launch) -> int {
    for (int w = 0; w < 3; w++) launch();
    CK(hipDeviceSynchronize());
    const int reps = 20;
    CK(hipEventRecord(e0));
    for (int r = 0; r < reps; r++) launch();
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("%-34s %8.1f GB/s  (%.3f ms per launch)\n", name, moved / (ms / reps * 1e-3) / 1e9, ms / reps);
    std::fflush(stdout);
    return 0;
  };
  const double cp = 2.0 * (double)bytes;
  if (rate("copy, 16 B per thread", cp, [&] { hipLaunchKernelGGL(k_copy16, dim3(g16), dim3(256), 0, 0, a, b, n16); }))
    return 1;
  if (rate("copy, 16 B per thread, nontemporal", cp,
           [&] { hipLaunchKernelGGL(k_copy16_nt, dim3(g16), dim3(256), 0, 0, a, b, n16); }))
    return 1;
  if (rate("copy, 8 B per thread", cp,
           [&] { hipLaunchKernelGGL(k_copy8, dim3(g8), dim3(256), 0, 0, (const double*)a, (double*)b, n8); }))
    return 1;
  if (rate("read only", (double)bytes, [&] { hipLaunchKernelGGL(k_read16, dim3(g16), dim3(256), 0, 0, a, o, n16); }))
    return 1;
  if (rate("write only", (double)bytes, [&] { hipLaunchKernelGGL(k_write16, dim3(g16), dim3(256), 0, 0, b, n16); }))
    return 1;
  if (rate("runtime copy (hipMemcpyAsync)", cp,
           [&] { (void)hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, 0); }))
    return 1;
  CK(hipFree(a));
  CK(hipFree(b));
  CK(hipFree(o));
  return 0;
}
