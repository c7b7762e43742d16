// VALU issue-rate probe: how many wave64 FP32 VALU instructions one SIMD
// issues per second when nothing else limits it.  Each lane runs 16
// independent chains of one instruction kind (v_add_f32: two source
// operands; v_fma_f32: three) for a fixed number of iterations, at eight
// waves per SIMD.  The density and force kernels' SQ_INSTS_VALU counts,
// divided by their durations, are compared with these rates.
//
// Prints one JSON line: wave-instructions per second per SIMD for each kind.
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

constexpr int kChains = 16;
constexpr int kThreads = 256;
constexpr int kIters = 10000;

template <bool kFma>
__global__ void __launch_bounds__(kThreads) chains(float* out, float z, int iters) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  float x[kChains], y[kChains];
#pragma unroll
  for (int c = 0; c < kChains; c++) {
    x[c] = gid * 1e-7f + c;
    y[c] = 0.5f + c * 0.01f + gid * 1e-9f;
  }
  for (int k = 0; k < iters; k++) {
#pragma unroll
    for (int c = 0; c < kChains; c++) x[c] = kFma ? fmaf(x[c], y[c], z) : x[c] + y[c];
  }
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < kChains; c++) s += x[c];
  out[gid] = s;  // one store per thread, inside the buffer of blocks * kThreads floats
}

bool ok(hipError_t e, const char* what) {
  if (e != hipSuccess) fprintf(stderr, "valu_rate: %s: %s\n", what, hipGetErrorString(e));
  return e == hipSuccess;
}

// Best of three timed launches, in wave-instructions per second per SIMD.
template <bool kFma>
bool rate(float* out, int blocks, int simds, double& r) {
  hipEvent_t e0, e1;
  if (!ok(hipEventCreate(&e0), "hipEventCreate") || !ok(hipEventCreate(&e1), "hipEventCreate")) return false;
  hipLaunchKernelGGL(chains<kFma>, dim3(blocks), dim3(kThreads), 0, 0, out, 0.999f, 10);
  if (!ok(hipDeviceSynchronize(), "warm-up launch")) return false;
  float best = 1e30f;
  for (int rep = 0; rep < 3; rep++) {
    if (!ok(hipEventRecord(e0), "hipEventRecord")) return false;
    hipLaunchKernelGGL(chains<kFma>, dim3(blocks), dim3(kThreads), 0, 0, out, 0.999f, kIters);
    if (!ok(hipEventRecord(e1), "hipEventRecord") || !ok(hipEventSynchronize(e1), "timed launch")) return false;
    float ms = 0.0f;
    if (!ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) return false;
    if (ms < best) best = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const double waves = (double)blocks * kThreads / 64.0;
  r = waves * kChains * kIters / (best * 1e-3) / simds;
  return true;
}

}  // namespace

int main() {
  int cus = 0;
  if (!ok(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0), "CU count") || cus <= 0) return 1;
  const int simds = 4 * cus;
  const int blocks = 8 * cus;  // 4 waves per block: 8 waves per SIMD
  float* out = nullptr;
  if (!ok(hipMalloc(&out, (size_t)blocks * kThreads * sizeof(float)), "hipMalloc")) return 1;
  double add = 0.0, fma = 0.0;
  const bool good = rate<false>(out, blocks, simds, add) && rate<true>(out, blocks, simds, fma);
  (void)hipFree(out);
  if (!good) return 1;
  printf("{\"simds\": %d, \"waves_per_simd\": 8, \"chains\": %d, \"add_f32_per_simd_s\": %.4g, "
         "\"fma_f32_per_simd_s\": %.4g}\n", simds, kChains, add, fma);
  return 0;
}
