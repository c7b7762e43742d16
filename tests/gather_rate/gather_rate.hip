// Scattered-gather rate probe: how many 16-byte lane loads per second the
// device serves when every lane of a wave reads its own 32-byte record (two
// 16-byte loads, as the force stage reads a neighbour's (pos, rho) and
// (veval, 0)) from a random place in a window of records, with three
// records' loads in flight per lane.  The force stage's pair gathers per
// second are compared with these rates.
//
// Windows: one 16 KB window that every block reads (it stays in each CU's
// 32 KB L1, so the loads are L1 hits: the rate of the vector-memory path
// itself for 64 distinct addresses per instruction), a 64 KB and a 256 KB
// window per block (L1 misses served by the L2), and the whole 128 MB buffer
// (HBM).  Prints one JSON line: lane loads per second for each window.
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

constexpr int kThreads = 256;
constexpr int kIters = 512;  // records per lane
constexpr int kAhead = 3;    // records whose loads are issued together
constexpr unsigned kRecords = 1u << 22;  // 4M records of 32 bytes

__global__ void __launch_bounds__(kThreads) gather(const float4* __restrict__ rec, float* out, unsigned wmask,
                                                   unsigned bmask, int iters) {
  const unsigned gid = blockIdx.x * blockDim.x + threadIdx.x;
  // the block's window: [base, base + wmask + 1) inside the buffer (base 0
  // for every block when bmask is 0)
  const unsigned base = (blockIdx.x * 4096u) & bmask & (kRecords - 1u) & ~wmask;
  unsigned x = gid * 2654435761u + 12345u;
  float s = 0.0f;
  for (int k = 0; k < iters; k += kAhead) {
    float4 a[kAhead], v[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; u++) {
      x ^= x << 13;
      x ^= x >> 17;
      x ^= x << 5;
      const unsigned j = base + (x & wmask);  // inside [0, kRecords): base + wmask < kRecords
      a[u] = rec[2 * (size_t)j];
      v[u] = rec[2 * (size_t)j + 1];
    }
#pragma unroll
    for (int u = 0; u < kAhead; u++) s += ((a[u].x + a[u].y) + (a[u].z + a[u].w)) + ((v[u].x + v[u].y) + (v[u].z + v[u].w));
  }
  out[gid] = s;  // one store per thread, inside the buffer of blocks * kThreads floats
}

bool ok(hipError_t e, const char* what) {
  if (e != hipSuccess) fprintf(stderr, "gather_rate: %s: %s\n", what, hipGetErrorString(e));
  return e == hipSuccess;
}

// Best of three timed launches, in 16-byte lane loads per second.
bool rate(const float4* rec, float* out, int blocks, unsigned wmask, unsigned bmask, double& r) {
  hipEvent_t e0, e1;
  if (!ok(hipEventCreate(&e0), "hipEventCreate") || !ok(hipEventCreate(&e1), "hipEventCreate")) return false;
  hipLaunchKernelGGL(gather, dim3(blocks), dim3(kThreads), 0, 0, rec, out, wmask, bmask, 3);
  if (!ok(hipDeviceSynchronize(), "warm-up launch")) return false;
  float best = 1e30f;
  for (int rep = 0; rep < 3; rep++) {
    if (!ok(hipEventRecord(e0), "hipEventRecord")) return false;
    hipLaunchKernelGGL(gather, dim3(blocks), dim3(kThreads), 0, 0, rec, out, wmask, bmask, kIters);
    if (!ok(hipEventRecord(e1), "hipEventRecord") || !ok(hipEventSynchronize(e1), "timed launch")) return false;
    float ms = 0.0f;
    if (!ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) return false;
    if (ms < best) best = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const int iters = (kIters + kAhead - 1) / kAhead * kAhead;
  r = (double)blocks * kThreads * iters * 2.0 / (best * 1e-3);
  return true;
}

}  // namespace

int main() {
  int cus = 0;
  if (!ok(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0), "CU count") || cus <= 0) return 1;
  const int blocks = 64 * cus;  // many rounds of 256-thread blocks
  float4* rec = nullptr;
  float* out = nullptr;
  if (!ok(hipMalloc(&rec, (size_t)kRecords * 2 * sizeof(float4)), "hipMalloc")) return 1;
  if (!ok(hipMalloc(&out, (size_t)blocks * kThreads * sizeof(float)), "hipMalloc")) return 1;
  if (!ok(hipMemset(rec, 0, (size_t)kRecords * 2 * sizeof(float4)), "hipMemset")) return 1;
  double r16 = 0.0, r64 = 0.0, r256 = 0.0, rall = 0.0;
  const bool good = rate(rec, out, blocks, (16u * 1024u / 32u) - 1u, 0u, r16) &&
                    rate(rec, out, blocks, (64u * 1024u / 32u) - 1u, ~0u, r64) &&
                    rate(rec, out, blocks, (256u * 1024u / 32u) - 1u, ~0u, r256) &&
                    rate(rec, out, blocks, kRecords - 1u, ~0u, rall);
  (void)hipFree(rec);
  (void)hipFree(out);
  if (!good) return 1;
  printf("{\"cus\": %d, \"threads\": %d, \"loads_in_flight_per_lane\": %d, \"lane_loads_per_s_16KB_shared\": %.4g, "
         "\"lane_loads_per_s_64KB\": %.4g, \"lane_loads_per_s_256KB\": %.4g, \"lane_loads_per_s_128MB\": %.4g}\n",
         cus, kThreads, 2 * kAhead, r16, r64, r256, rall);
  return 0;
}
