// refhost.cpp -- a host that binds particles_gfx950.hsaco the way the
// reference's Fluids 5.0 host binds its own particles.ptx, as a drop-in check
// of the module contract.  Test infrastructure: run by tests/test_gpu_refhost.py.
//
// What it does, as the reference does it:
//   - loads the code object and the seven kernels by name, and the four
//     module globals FParams, FPnts, FPntTmp, FAccel with their size checks
//     (particles.cpp:68-93, datax.cpp:68-80);
//   - allocates the eight particle buffers of ReallocateParticles (no id
//     buffer) twice, points and temp (particles.cpp:267-300), and the seven
//     acceleration buffers with gridTotal-entry count and offset tables
//     (particles.cpp:596-612); unused table slots stay 0 (datax.cpp:81-94);
//   - launches with 512 threads per block (particles.cpp:26-30,294-296) on the
//     NULL stream, with the per-step memsets of InsertParticles
//     (particles.cpp:721-725), the scan launches of PrefixScanParticles
//     (:774-815), the eight device copies of CopyAllToTemp before
//     countingSortFull (:698-701,832-842), then computePressure,
//     computeForce and advanceParticles(time, dt, ss, pnum) and a
//     synchronize per step (:382-394,413-418,852-857,914-919).
// FHostCaps is never written, so the module runs as it would under the
// reference's host.
//
//   refhost <code object> <in.bin> <out.bin> <nsteps>
//   in.bin : FParams_t (688 B) | int32 n | float pos[3n] | uint32 clr[n]
//   out.bin: int32 n | int32 G | the eight FPnts buffers in id order 0..7
//            | AGRIDCNT (G) | AGRIDOFF (G) | AGRID (n)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fluidsx_params.h"

static void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    fprintf(stderr, "refhost: %s: %s\n", what, hipGetErrorString(e));
    exit(1);
  }
}

namespace {

const char* kNames[7] = {"insertParticles", "countingSortFull", "computePressure", "computeForce",
                         "advanceParticles", "prefixSum", "prefixFixup"};
enum { K_INSERT, K_SORT, K_PRESS, K_FORCE, K_ADVANCE, K_SCAN, K_FIXUP };

// strides of the reference's eight particle buffers by id (fluid.h:15-22)
const size_t kStride[8] = {12, 12, 4, 12, 12, 4, 4, 4};

struct Table {  // one DataX: device pointers by id and the module global they are published to
  hipDeviceptr_t sym = nullptr;
  uint64_t mbuf[FX_DMAXBUF] = {};
  void publish() { check(hipMemcpyHtoD(sym, mbuf, sizeof(mbuf)), "UpdateGPUAccess"); }
};

void launch(hipFunction_t f, int blocks, int threads, void** args, const char* what) {
  check(hipModuleLaunchKernel(f, blocks, 1, 1, threads, 1, 1, 0, nullptr, args, nullptr), what);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: refhost <code object> <in.bin> <out.bin> <nsteps>\n");
    return 2;
  }
  const int nsteps = atoi(argv[4]);
  FParams_t P;
  int n = 0;
  FILE* fi = fopen(argv[2], "rb");
  if (!fi || fread(&P, sizeof(P), 1, fi) != 1 || fread(&n, 4, 1, fi) != 1 || n <= 0) {
    fprintf(stderr, "refhost: cannot read %s\n", argv[2]);
    return 1;
  }
  std::vector<float> pos((size_t)3 * n);
  std::vector<uint32_t> clr(n);
  if (fread(pos.data(), 4, pos.size(), fi) != pos.size() || fread(clr.data(), 4, n, fi) != (size_t)n) {
    fprintf(stderr, "refhost: short input\n");
    return 1;
  }
  fclose(fi);

  // Initialize
  check(hipInit(0), "hipInit");
  hipModule_t mod;
  check(hipModuleLoad(&mod, argv[1]), "hipModuleLoad");
  hipFunction_t fn[7];
  for (int k = 0; k < 7; k++) check(hipModuleGetFunction(&fn[k], mod, kNames[k]), kNames[k]);
  Table pnts, ptmp, accel;
  size_t len = 0;
  const char* gname[3] = {"FPnts", "FPntTmp", "FAccel"};
  Table* tabs[3] = {&pnts, &ptmp, &accel};
  for (int k = 0; k < 3; k++) {
    check(hipModuleGetGlobal(&tabs[k]->sym, &len, mod, gname[k]), gname[k]);
    if (len != sizeof(cuDataX)) {
      fprintf(stderr, "refhost: %s has %zu bytes, cuDataX %zu\n", gname[k], len, sizeof(cuDataX));
      return 1;
    }
  }
  hipDeviceptr_t cuParams;
  check(hipModuleGetGlobal(&cuParams, &len, mod, "FParams"), "FParams");
  if (len != sizeof(FParams_t)) {
    fprintf(stderr, "refhost: FParams has %zu bytes\n", len);
    return 1;
  }

  // ReallocateParticles: eight buffers, points and temp, 512 threads per block
  for (int b = 0; b < 8; b++) {
    void* d = nullptr;
    check(hipMalloc(&d, kStride[b] * n), "cuMemAlloc points");
    pnts.mbuf[b] = (uint64_t)d;
    check(hipMalloc(&d, kStride[b] * n), "cuMemAlloc temp");
    ptmp.mbuf[b] = (uint64_t)d;
  }
  const int threads = n < 512 ? n : 512;
  const int blocks = (n + threads - 1) / threads;
  P.pnum = n;
  P.numThreads = threads;
  P.numBlocks = blocks;
  P.szPnts = threads * blocks;
  pnts.publish();
  ptmp.publish();

  // RebuildAccelGrid: gridTotal-entry count and offset tables
  const int G = P.gridTotal;
  const int elem1 = G, elem2 = elem1 / 1024 + 1, elem3 = elem2 / 1024 + 1;
  const size_t asz[7] = {(size_t)n, (size_t)G, (size_t)G, (size_t)elem2, (size_t)elem2, (size_t)elem3, (size_t)elem3};
  for (int b = 0; b < 7; b++) {
    void* d = nullptr;
    check(hipMalloc(&d, 4 * asz[b]), "cuMemAlloc accel");
    accel.mbuf[b] = (uint64_t)d;
  }
  accel.publish();
  check(hipMemcpyHtoD(cuParams, &P, sizeof(P)), "UpdateParams");

  // CommitAll: positions and colours, every other field zero
  std::vector<char> zero((size_t)12 * n, 0);
  for (int b = 0; b < 8; b++) {
    const void* src = b == FPOS ? (const void*)pos.data() : b == FCLR ? (const void*)clr.data() : zero.data();
    check(hipMemcpyHtoD((hipDeviceptr_t)pnts.mbuf[b], const_cast<void*>(src), kStride[b] * n), "CommitAll");
  }

  auto dp = [](uint64_t a) { return (hipDeviceptr_t)a; };
  for (int step = 0; step < nsteps; step++) {
    // InsertParticles
    check(hipMemsetD8(dp(accel.mbuf[AGRIDCNT]), 0, 4 * (size_t)G), "memset AGRIDCNT");
    check(hipMemsetD8(dp(accel.mbuf[AGRIDOFF]), 0, 4 * (size_t)G), "memset AGRIDOFF");
    check(hipMemsetD8(dp(pnts.mbuf[FGCELL]), 0, 4 * (size_t)n), "memset FGCELL");
    check(hipMemsetD8(dp(pnts.mbuf[FGNDX]), 0, 4 * (size_t)n), "memset FGNDX");
    void* ia[] = {&n};
    launch(fn[K_INSERT], blocks, threads, ia, "insertParticles");
    // PrefixScanParticles
    {
      int e1 = elem1, e2 = elem2, e3 = elem3, zon = 1;
      uint64_t a1 = accel.mbuf[AGRIDCNT], s1 = accel.mbuf[AGRIDOFF], a2 = accel.mbuf[AAUXARRAY1],
               s2 = accel.mbuf[AAUXSCAN1], a3 = accel.mbuf[AAUXARRAY2], s3 = accel.mbuf[AAUXSCAN2], none = 0;
      void* A[] = {&a1, &s1, &a2, &e1, &zon};
      launch(fn[K_SCAN], e2, 512, A, "prefixSum A");
      void* B[] = {&a2, &s2, &a3, &e2, &zon};
      launch(fn[K_SCAN], e3, 512, B, "prefixSum B");
      if (e3 > 1) {
        void* Cc[] = {&a3, &s3, &none, &e3, &zon};
        launch(fn[K_SCAN], 1, 512, Cc, "prefixSum C");
        void* D[] = {&s2, &s3, &e2};
        launch(fn[K_FIXUP], e3, 512, D, "prefixFixup D");
      }
      void* E[] = {&s1, &s2, &e1};
      launch(fn[K_FIXUP], e2, 512, E, "prefixFixup E");
    }
    // CountingSort: CopyAllToTemp, then the scatter
    for (int b = 0; b < 8; b++)
      check(hipMemcpyDtoD(dp(ptmp.mbuf[b]), dp(pnts.mbuf[b]), kStride[b] * n), "CopyAllToTemp");
    launch(fn[K_SORT], blocks, threads, ia, "countingSortFull");
    launch(fn[K_PRESS], blocks, threads, ia, "computePressure");
    launch(fn[K_FORCE], blocks, threads, ia, "computeForce");
    float time = P.time, dt = P.dt, ss = P.sim_scale;
    int pnum = P.pnum;
    void* aa[] = {&time, &dt, &ss, &pnum};
    launch(fn[K_ADVANCE], blocks, threads, aa, "advanceParticles");
    check(hipDeviceSynchronize(), "cuCtxSynchronize");
    P.time += P.dt;  // AdvanceTime (the next UpdateParams is not needed: kernels take time as an argument)
  }

  // results
  FILE* fo = fopen(argv[3], "wb");
  if (!fo) {
    fprintf(stderr, "refhost: cannot write %s\n", argv[3]);
    return 1;
  }
  fwrite(&n, 4, 1, fo);
  fwrite(&G, 4, 1, fo);
  std::vector<char> host((size_t)12 * n);
  for (int b = 0; b < 8; b++) {
    check(hipMemcpyDtoH(host.data(), dp(pnts.mbuf[b]), kStride[b] * n), "Retrieve");
    fwrite(host.data(), 1, kStride[b] * n, fo);
  }
  const int abuf[3] = {AGRIDCNT, AGRIDOFF, AGRID};
  for (int k = 0; k < 3; k++) {
    const size_t cnt = k == 2 ? (size_t)n : (size_t)G;
    check(hipMemcpyDtoH(host.data(), dp(accel.mbuf[abuf[k]]), 4 * cnt), "Retrieve accel");
    fwrite(host.data(), 4, cnt, fo);
  }
  fclose(fo);
  for (int b = 0; b < 8; b++) {
    (void)hipFree((void*)pnts.mbuf[b]);
    (void)hipFree((void*)ptmp.mbuf[b]);
  }
  for (int b = 0; b < 7; b++) (void)hipFree((void*)accel.mbuf[b]);
  (void)hipModuleUnload(mod);
  printf("refhost: %d particles, %d cells, %d steps\n", n, G, nsteps);
  return 0;
}
