// issue_probe.hip -- VALU issue rates of this GPU: independent FMA chains in
// FP32 and FP64 (8 chains and 1 chain per thread) and a v_rcp_f32 + add
// chain, timed with HIP events; prints wave-instructions per SIMD-cycle
// (1,024 SIMDs at 2.4 GHz).  Built by awflcloud_amd/build.py as
// awflcloud_amd/issue_probe; every thread writes one value to out[its id],
// and out holds blocks*256 doubles.
//   awflcloud_amd/issue_probe
#include <hip/hip_runtime.h>
#include <cstdio>
#include <vector>
#define CHK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)
template <class T, int NCH>
__global__ __launch_bounds__(256) void k_fma(T* out, T a, T b, int iters) {
  T x[NCH];
#pragma unroll
  for (int c = 0; c < NCH; c++) x[c] = (T)(threadIdx.x + c);
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int r = 0; r < 16; r++) {
#pragma unroll
      for (int c = 0; c < NCH; c++) x[c] = x[c] * a + b;
    }
  }
  T s = 0;
#pragma unroll
  for (int c = 0; c < NCH; c++) s += x[c];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
template <int NCH>
__global__ __launch_bounds__(256) void k_rcp(float* out, float a, int iters) {
  float x[NCH];
#pragma unroll
  for (int c = 0; c < NCH; c++) x[c] = 1.0f + threadIdx.x + c;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int r = 0; r < 16; r++) {
#pragma unroll
      for (int c = 0; c < NCH; c++) x[c] = __builtin_amdgcn_rcpf(x[c]) + a;
    }
  }
  float s = 0;
#pragma unroll
  for (int c = 0; c < NCH; c++) s += x[c];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
int main() {
  const int blocks = 256 * 8 * 4;   // 8 waves per SIMD-ish
  const int iters = 2000;
  void* d;
  CHK(hipMalloc(&d, sizeof(double) * blocks * 256));
  hipEvent_t e0, e1;
  CHK(hipEventCreate(&e0)); CHK(hipEventCreate(&e1));
  auto time = [&](auto launch, const char* name, double instr_per_thread) -> int {
    launch();
    CHK(hipDeviceSynchronize());
    CHK(hipEventRecord(e0));
    launch();
    CHK(hipEventRecord(e1));
    CHK(hipEventSynchronize(e1));
    float ms;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    double waveinstr = instr_per_thread * blocks * 256 / 64.0;
    double per_simd_cycle = waveinstr / (1024.0 * ms * 1e-3 * 2.4e9);
    printf("%-12s %8.3f ms  wave-instr per SIMD-cycle %.3f  (cycles per wave64 instr %.2f)\n", name, ms, per_simd_cycle, 1.0 / per_simd_cycle);
    return 0;
  };
  double n = 16.0 * iters;
  if (time([&] { k_fma<float, 8><<<blocks, 256>>>((float*)d, 0.999f, 0.5f, iters); }, "fma_f32 x8", 8 * n)) return 1;
  if (time([&] { k_fma<float, 4><<<blocks, 256>>>((float*)d, 0.999f, 0.5f, iters); }, "fma_f32 x4", 4 * n)) return 1;
  if (time([&] { k_fma<float, 1><<<blocks, 256>>>((float*)d, 0.999f, 0.5f, iters); }, "fma_f32 x1", 1 * n)) return 1;
  if (time([&] { k_fma<double, 8><<<blocks, 256>>>((double*)d, 0.999, 0.5, iters); }, "fma_f64 x8", 8 * n)) return 1;
  if (time([&] { k_fma<double, 1><<<blocks, 256>>>((double*)d, 0.999, 0.5, iters); }, "fma_f64 x1", 1 * n)) return 1;
  if (time([&] { k_rcp<8><<<blocks, 256>>>((float*)d, 0.5f, iters); }, "rcp+add f32", 16 * n)) return 1;
  CHK(hipFree(d));
  return 0;
}
