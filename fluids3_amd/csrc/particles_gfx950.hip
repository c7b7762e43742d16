// particles_gfx950.hip -- the SPH step as hand-written gfx950 (CDNA4, wave64)
// kernels, built into one code object that keeps the reference's module
// contract: the same extern "C" kernel names and argument lists
// (fluids5.0/src/particles.cuh:29-42) and the same four __constant__ symbols
// (fluids5.0/src/particles.cu:17-20), so a host that loads kernels and
// globals by name can drive it.
//
// Every stage is a device function over one scene: its FParams_t and its
// particle and cell ranges.  The reference-named kernels run it for the
// single scene held in FParams; the *Batch kernels run it for many
// independent scenes concatenated in the same buffers, one launch per stage,
// each block working inside one scene (a table maps block -> scene).
//
// Arithmetic contract (shared with the CPU oracle, which evaluates the same
// expression trees): the file is compiled with -ffp-contract=off, every fused
// multiply-add is an explicit fmaf(), sqrtf and '/' are IEEE correctly
// rounded.  With the stable in-cell order produced by the counting sort the
// FP32 results are then reproducible run to run and comparable bit for bit.
#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>

#include "fluidsx_params.h"
#include "fx_trig.hpp"

typedef unsigned int uint;

__constant__ FParams_t FParams;
__constant__ cuDataX FPnts;
__constant__ cuDataX FPntTmp;
__constant__ cuDataX FAccel;

#define SCAN_BLOCKSIZE 512  // threads per scan block; each covers 1024 cells

// Words of neighbour bitmask a particle may record: FX_NBR_WORDS unless the
// host lowers it while the bitmask buffer is a smaller one (a reverse-mode
// tape slot).  A particle whose stencil rows need more records only the
// first ones, which the force stage does not read: it tests that particle's
// candidates in place.
__constant__ int FNbrWords = FX_NBR_WORDS;
__device__ __forceinline__ int nbr_cap() { return min(FNbrWords, FX_NBR_WORDS); }

// FX_CAP_* bits of what the host set up beyond the reference's layout; 0 for
// a host that binds this module as Fluids 5.0 binds its own.
__constant__ int FHostCaps = 0;
__device__ __forceinline__ bool host_tail() { return (FHostCaps & FX_CAP_TAIL) != 0; }
__device__ __forceinline__ bool host_ranked() { return (FHostCaps & FX_CAP_RANKED) != 0; }

// When set (the reverse mode's rollouts), the ranked counting sort writes
// its permutation here: FSortSrc[s] = pre-sort index of the particle it puts
// in slot s.  Null otherwise.
__constant__ uint* FSortSrc = nullptr;
#define FX_RANK_CAP 65536   // cells with more members keep atomic order

// ---------------------------------------------------------------- helpers

// Buffer id -> pointer.  The table holds global-memory addresses; passing
// them through the global address space lets the compiler emit global_*
// instructions instead of flat ones (flat accesses also count against the
// LDS wait counter and force coarser waits).
template <class T>
__device__ __forceinline__ T* pbuf(const cuDataX& d, int id) {
  typedef __attribute__((address_space(1))) T gT;
  return (T*)(gT*)(d.mbuf[id]);
}

// A scene's parameters read through the constant address space, so that the
// (block-uniform) reads are scalar loads.
// Three floats of element i, loaded and stored as scalars: an fx_float3
// copied as a whole (a struct select, say) can leave the compiler an
// aggregate it places in LDS or scratch memory instead of registers.
__device__ __forceinline__ fx_float3 ld3(const fx_float3* __restrict__ p, size_t i) {
  const float* f = reinterpret_cast<const float*>(p) + 3 * i;
  return fx_float3{f[0], f[1], f[2]};
}
// ld3 with non-temporal loads, for data read once
__device__ __forceinline__ fx_float3 ld3nt(const fx_float3* __restrict__ p, size_t i) {
  const float* f = reinterpret_cast<const float*>(p) + 3 * i;
  return fx_float3{__builtin_nontemporal_load(f), __builtin_nontemporal_load(f + 1), __builtin_nontemporal_load(f + 2)};
}
__device__ __forceinline__ void st3(fx_float3* __restrict__ p, size_t i, float x, float y, float z) {
  float* f = reinterpret_cast<float*>(p) + 3 * i;
  f[0] = x;
  f[1] = y;
  f[2] = z;
}

__device__ __forceinline__ const FParams_t* cparams(const FParams_t* p) {
  typedef const __attribute__((address_space(4))) FParams_t cP;
  return (const FParams_t*)(cP*)p;
}

// One block's view of its scene: particles [p0, p1), first cell c0 of the
// scene's cell table, and the block's index inside the scene.
struct Span {
  int p0, p1, c0, blk;
};

// Batch lookup: entry b of the block table is (scene, block inside scene).
// With a queue the launch is persistent: a fixed grid pulls the nitems
// entries as work items (see batch_items).
struct BatchArgs {
  const FParams_t* params;  // one FParams_t per scene
  const int* pbase;         // first particle of each scene (nscene + 1 entries)
  const int* cbase;         // first cell of each scene's table (nscene + 1)
  const int2* btab;         // (scene, local block) per block-table entry
  uint* queue;              // FX_NUM_XCD claim counters, zero at launch; null: one block per entry
  int nitems;               // entries of btab
  int pad;
};

__device__ __forceinline__ Span batch_span(const BatchArgs& B, int b, const FParams_t*& P) {
  const int2 e = B.btab[b];
  P = cparams(B.params + e.x);
  return Span{B.pbase[e.x], B.pbase[e.x + 1], B.cbase[e.x], e.y};
}

// Inclusive scan of one uint per lane across a wave64 with DPP row shifts and
// row broadcasts (GFX9 family).  Lanes whose DPP source is outside the row
// receive 0 through the 'old' operand.
__device__ __forceinline__ uint wave_inclusive_scan(uint v) {
  uint t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  v += t;
  return v;
}

// Wave-wide max / min with the same DPP pattern; lane 63 ends up with the
// reduction of all 64 lanes and is read back as a scalar.
template <bool kMax>
__device__ __forceinline__ int wave_reduce(int v) {
  const int id = kMax ? INT_MIN : INT_MAX;
  int t;
#define FX_RED_STEP(ctrl, rmask)                                    \
  t = __builtin_amdgcn_update_dpp(id, v, ctrl, rmask, 0xf, false); \
  v = kMax ? max(v, t) : min(v, t);
  FX_RED_STEP(0x111, 0xf)  // row_shr:1
  FX_RED_STEP(0x112, 0xf)  // row_shr:2
  FX_RED_STEP(0x114, 0xf)  // row_shr:4
  FX_RED_STEP(0x118, 0xf)  // row_shr:8
  FX_RED_STEP(0x142, 0xa)  // row_bcast:15
  FX_RED_STEP(0x143, 0xc)  // row_bcast:31
#undef FX_RED_STEP
  return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_max(int v) { return wave_reduce<true>(v); }
__device__ __forceinline__ int wave_min(int v) { return wave_reduce<false>(v); }

// Block barrier that orders LDS accesses only: global loads and atomics
// already in flight stay in flight across it (a __syncthreads() would first
// wait for them).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ uint lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((uint)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint)mask, 0u));
}

// XCD-aware block order.  MI355X hands consecutive workgroups to its eight
// XCDs round-robin, and each XCD has its own L2.  Cell-sorted particles that
// are close in index share neighbours, so runs of FX_XCD_RUN consecutive
// logical blocks are given to one XCD, and the runs rotate over the XCDs (a
// contiguous share per XCD would unbalance the load, since the work per
// particle follows the local density).  Dispatch slot b is the k-th block
// of XCD x = b % 8; it runs logical block (run * 8 + x) * FX_XCD_RUN + k % RUN
// with run = k / RUN.  Blocks past the last full round keep their index.
#define FX_NUM_XCD 8
#ifndef FX_XCD_RUN
#define FX_XCD_RUN 32
#endif
__device__ __forceinline__ int xcd_block() {
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int round = FX_NUM_XCD * FX_XCD_RUN;
  if (b >= nb / round * round) return b;
  const int x = b % FX_NUM_XCD, k = b / FX_NUM_XCD;
  return ((k / FX_XCD_RUN) * FX_NUM_XCD + x) * FX_XCD_RUN + k % FX_XCD_RUN;
}

// The XCD this workgroup runs on (hardware register XCC_ID, bits 3:0).
__device__ __forceinline__ int xcc_id() { return (int)(__builtin_amdgcn_s_getreg(20 | (3 << 11)) & 7u); }

// Persistent launches.  A fixed grid of resident blocks pulls work items
// (entries of the block table) until none is left.  The items are split into
// FX_NUM_XCD contiguous ranges, one per XCD, each drained in order through
// its own counter by the blocks running on that XCD, so that the blocks an
// XCD holds at a time work on neighbouring particles, whose neighbour cells
// then stay in that XCD's L2; a block whose range is exhausted takes items
// from the following ranges in turn.  A block leaves once every range is
// exhausted, so the launch drains whatever the grid size.  Thread 0 claims,
// the block reads the item from LDS.
// The claim state (current range, ranges tried) lives in LDS beside the
// claimed item: only thread 0 uses it, and it stays out of the registers of
// the work the block runs between claims.
struct Claim {
  int item, q, tried;
};

__device__ __forceinline__ int next_item(uint* ctr, int n, Claim* c) {
  if (threadIdx.x == 0) {
    int it = -1;
    int q = c->q, tried = c->tried;
    while (tried < FX_NUM_XCD) {
      const int lo = (int)(((long long)n * q) / FX_NUM_XCD);
      const int hi = (int)(((long long)n * (q + 1)) / FX_NUM_XCD);
      const int k = (int)atomicAdd(&ctr[q], 1u);
      if (k < hi - lo) {
        it = lo + k;
        break;
      }
      q = (q + 1) % FX_NUM_XCD;
      tried++;
    }
    c->item = it;
    c->q = q;
    c->tried = tried;
  }
  __syncthreads();
  const int it = c->item;
  __syncthreads();  // every thread has read the item before thread 0 claims the next
  return it;
}

// Runs body(entry) for every block-table entry this block claims in a
// persistent launch.
template <class F>
__device__ __forceinline__ void batch_items(const BatchArgs& B, F&& body) {
  __shared__ Claim c;
  if (threadIdx.x == 0) c = Claim{-1, xcc_id(), 0};
  for (int it = next_item(B.queue, B.nitems, &c); it >= 0; it = next_item(B.queue, B.nitems, &c)) body(it);
}

// Particles per thread of the streaming stages (index: FX_STAGE_INSERT 0,
// FX_STAGE_SORT 2, FX_STAGE_ADVANCE 5): a block covers blockDim * ppt
// consecutive particles, thread t taking t, t + blockDim, ...  A host that
// does not set it runs one particle per thread, as the reference's launches
// do (particles.cpp:26-30,294-296).
#define FX_MAX_PPT 8
__constant__ int FPpt[8] = {1, 1, 1, 1, 1, 1, 1, 1};
__device__ __forceinline__ int ppt_of(int stage) { return max(1, min(FPpt[stage], FX_MAX_PPT)); }
// the k-th of a block's ppt sub-blocks of blockDim particles
__device__ __forceinline__ Span sub_span(Span S, int ppt, int k) {
  S.blk = S.blk * ppt + k;
  return S;
}

// ----------------------------------------------------------- fill (memset)

// Fill used instead of hipMemsetAsync so that a captured step graph holds
// kernel nodes only.
extern "C" __global__ void fxFillUint(uint* __restrict__ p, uint v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// No work: one dispatch of it marks a point of the stream in kernel traces
// and counter runs (the benchmark brackets its timed region with two).
extern "C" __global__ void fxMark(int tag) {}

// ------------------------------------------------------------------ insert

// Cell of each particle and its count slot.  Follows particles.cu:24-54:
//   gc = int((p - gridMin) * gridDelta), gs = (gc.y*res.z + gc.z)*res.x + gc.x,
//   in range iff 1 <= gc <= gridScanMax on every axis, else GRID_UNDEF.
// The range test is done on the float before truncation (trunc(g) >= 1 iff
// g >= 1; trunc(g) <= M iff g < M+1), which is the same predicate for every
// finite input and sends NaN to GRID_UNDEF.  Out-of-range particles are
// counted in slot gridTotal of the scene (its tail bucket).  Stored cell
// indices are offsets into the concatenated cell table (c0 + local cell).
// Without a tail slot (tail = false: a table of gridTotal entries, as the
// reference allocates) out-of-range particles are not counted and keep the
// gndx the host cleared, as in particles.cu:45-52.
//
// The counting is aggregated per wave: the particles arrive sorted by cell
// from the previous step, so the 64 lanes of a wave fall into few cells;
// lanes of one cell take a single atomicAdd of their number and their slot
// is the returned base plus their rank among those lanes.  The atomics of
// all the wave's cells are issued together (no round trip per cell).  The
// tail bucket is aggregated per block instead: particles that leave the grid
// stay out of it (the advance parks them), they are sorted to the end, and
// in a long run whole blocks hold nothing else (a sixth of the particles by
// step 200 of the 4M dam break), so one count per wave would put thousands
// of atomics on that one address.  `tl` is the block's LDS (two sets of 17
// words, alternating between calls so that one call's reads cannot meet
// the next call's writes).
__device__ __forceinline__ void insert_impl(const FParams_t& P, const Span S, bool tail, uint (*tl)[17]) {
  const int i = S.p0 + S.blk * (int)blockDim.x + (int)threadIdx.x;
  bool active = i < S.p1;  // below: particles that take a count slot
  int gs = 0;
  if (active) {
    const fx_float3 p = pbuf<const fx_float3>(FPnts, FPOS)[i];
    const float gx = (p.x - P.gridMin.x) * P.gridDelta.x;
    const float gy = (p.y - P.gridMin.y) * P.gridDelta.y;
    const float gz = (p.z - P.gridMin.z) * P.gridDelta.z;
    const bool in = gx >= 1.0f && gx < (float)(P.gridScanMax.x + 1) && gy >= 1.0f &&
                    gy < (float)(P.gridScanMax.y + 1) && gz >= 1.0f && gz < (float)(P.gridScanMax.z + 1);
    if (in) {
      gs = S.c0 + ((int)gy * P.gridRes.z + (int)gz) * P.gridRes.x + (int)gx;
      pbuf<int>(FPnts, FGCELL)[i] = gs;
    } else {
      gs = S.c0 + P.gridTotal;
      pbuf<int>(FPnts, FGCELL)[i] = FX_GRID_UNDEF;
      active = tail;
    }
    uint* tflag = pbuf<uint>(FAccel, ATAILFLAG);
    if (tflag != nullptr) tflag[i] = in ? 0u : 1u;
  }
  uint* cnt = pbuf<uint>(FAccel, AGRIDCNT);
  const int tslot = S.c0 + P.gridTotal;
  const bool intail = active && gs == tslot;
  // group the wave's lanes by slot (register-only ballots), then every
  // group's first lane adds the group's size with one atomic; all groups'
  // atomics go out together and each lane reads its group's base from the
  // leader lane
  unsigned long long todo = __ballot(active && !intail);
  unsigned long long mine = 0ull;
  int leader_of = 0;
  while (todo != 0ull) {  // one pass per distinct slot in the wave
    const int leader = __ffsll((long long)todo) - 1;
    const int v = __builtin_amdgcn_readlane(gs, leader);
    const unsigned long long same = __ballot(active && gs == v) & todo;
    if (active && gs == v) {
      mine = same;
      leader_of = leader;
    }
    todo &= ~same;
  }
  uint base = 0, tbase = 0;
  if (active && !intail && (int)__lane_id() == leader_of) base = atomicAdd(&cnt[gs], (uint)__popcll(mine));
  unsigned long long tm = 0ull;
  if (tail) {  // uniform: the block's tail lanes, one atomic (the cell atomics stay in flight meanwhile)
    tm = __ballot(intail);
    const int wv = (int)(threadIdx.x >> 6), nw = (int)((blockDim.x + 63u) >> 6);
    uint* L = tl[S.blk & 1];
    if (__lane_id() == 0) L[wv] = (uint)__popcll(tm);
    lds_barrier();
    if (threadIdx.x == 0) {
      uint tot = 0;
      for (int w = 0; w < nw; w++) tot += L[w];
      L[16] = tot != 0u ? atomicAdd(&cnt[tslot], tot) : 0u;
    }
    lds_barrier();
    if (intail) {
      tbase = L[16];
      for (int w = 0; w < wv; w++) tbase += L[w];
    }
  }
  base = (uint)__shfl((int)base, leader_of, 64);
  if (intail) {
    base = tbase;
    mine = tm;
  }
  if (active) pbuf<int>(FPnts, FGNDX)[i] = (int)(base + lanes_below(mine));
}

// The streaming kernels keep a straight-line body for one particle per
// thread (the reference's launches and the usual choice), the loop otherwise.
extern "C" __global__ void insertParticles(int pnum) {
  __shared__ uint tl[2][17];
  const int ppt = ppt_of(0);
  const Span S{0, pnum, 0, (int)blockIdx.x};
  if (ppt == 1) {
    insert_impl(*cparams(&FParams), S, host_tail(), tl);
    return;
  }
  for (int k = 0; k < ppt; k++) insert_impl(*cparams(&FParams), sub_span(S, ppt, k), host_tail(), tl);
}

// one block-table entry of a batch stage (the *Batch kernels run one per
// block, the *BatchP kernels claim them)
__device__ __forceinline__ void insert_entry(const BatchArgs& B, int b) {
  const int ppt = ppt_of(0);
  const FParams_t* P;
  const Span S = batch_span(B, b, P);
  __shared__ uint tl[2][17];
  for (int k = 0; k < ppt; k++) insert_impl(*P, sub_span(S, ppt, k), true, tl);
}

extern "C" __global__ void insertParticlesBatch(BatchArgs B) { insert_entry(B, blockIdx.x); }

extern "C" __global__ void insertParticlesBatchP(BatchArgs B) {
  batch_items(B, [&](int b) { insert_entry(B, b); });
}

// ------------------------------------------------------------- prefix scan

// Block-local scan of 1024 uints, reference semantics (particles.cu:333-369):
// block b covers [1024b, 1024b+1024); zeroff=1 gives the exclusive scan,
// zeroff=0 the inclusive one; aux[b] (if given) receives the block total.
// Each thread owns two adjacent elements; waves scan with DPP, the eight wave
// totals are combined through LDS.  A batch of scenes is scanned in one pass
// over the concatenated cell tables: every scene's counts (tail bucket
// included) add up to its particle count, so the running sum enters each
// scene at its first particle.
extern "C" __global__ __launch_bounds__(SCAN_BLOCKSIZE) void prefixSum(
    uint* input, uint* output, uint* aux, int len, int zeroff) {
  __shared__ uint wsum[SCAN_BLOCKSIZE / 64];
  const int t = threadIdx.x;
  const uint e0 = blockIdx.x * (2u * SCAN_BLOCKSIZE) + 2u * t;
  const uint e1 = e0 + 1;
  const uint ulen = (uint)len;
  const uint a = e0 < ulen ? input[e0] : 0u;
  const uint b = e1 < ulen ? input[e1] : 0u;
  const uint s = a + b;
  const uint inc = wave_inclusive_scan(s);
  const int lane = t & 63, wave = t >> 6;
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  uint wpre = 0, total = 0;
#pragma unroll
  for (int w = 0; w < SCAN_BLOCKSIZE / 64; w++) {
    const uint v = wsum[w];
    wpre += (w < wave) ? v : 0u;
    total += v;
  }
  const uint excl = wpre + inc - s;
  if (zeroff) {
    if (e0 < ulen) output[e0] = excl;
    if (e1 < ulen) output[e1] = excl + a;
  } else {
    if (e0 < ulen) output[e0] = excl + a;
    if (e1 < ulen) output[e1] = excl + a + b;
  }
  if (aux != nullptr && t == 0) aux[blockIdx.x] = total;
}

// Adds aux[b] to the 1024 elements of block b (particles.cu:325-331).
extern "C" __global__ __launch_bounds__(SCAN_BLOCKSIZE) void prefixFixup(
    uint* input, uint* aux, int len) {
  const uint t = threadIdx.x;
  const uint start = t + 2u * blockIdx.x * SCAN_BLOCKSIZE;
  const uint add = aux[blockIdx.x];
  if (start < (uint)len) input[start] += add;
  if (start + SCAN_BLOCKSIZE < (uint)len) input[start + SCAN_BLOCKSIZE] += add;
}

// The step's own scans (cell counts -> offsets, out-of-range flags -> tail
// ranks) in two launches each instead of the three to five of the
// reference's prefixSum/prefixFixup levels: fxScanSums writes the total of
// every block of FXS_BLOCK elements; in fxScanApply block b adds up the
// totals of the blocks before it (a few thousand at most, read once per
// block) and scans its own elements on top.  Exclusive, uint32: the same
// results as the reference-named kernels.
#define FXS_THREADS 512
#define FXS_PER 4
#define FXS_BLOCK (FXS_THREADS * FXS_PER)

// elements i..i+3 (zero beyond n); one 16-byte load when all four exist
__device__ __forceinline__ uint4 fxs_load(const uint* __restrict__ in, int i, int n) {
  if (i + 3 < n) return *reinterpret_cast<const uint4*>(in + i);
  uint4 v = make_uint4(0u, 0u, 0u, 0u);
  if (i < n) v.x = in[i];
  if (i + 1 < n) v.y = in[i + 1];
  if (i + 2 < n) v.z = in[i + 2];
  return v;
}

// sum of v over the block, returned to every thread (lds: one uint per wave)
__device__ __forceinline__ uint fxs_block_sum(uint v, uint* lds) {
  const uint tot = (uint)__builtin_amdgcn_readlane((int)wave_inclusive_scan(v), 63);
  if ((threadIdx.x & 63u) == 0u) lds[threadIdx.x >> 6] = tot;
  __syncthreads();
  uint s = 0u;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += lds[w];
  __syncthreads();  // every thread has read lds before it is written again
  return s;
}

// One launch serves up to two independent scans (the step scans the cell
// counts and the out-of-range flags): blocks [0, j[0].nb) take the first,
// the rest the second.
struct FxScanJob {
  const uint* in;
  uint* out;
  uint* sums;  // one total per block of FXS_BLOCK elements
  int n;
  int nb;  // blocks: ceil(n / FXS_BLOCK)
};
struct FxScanJobs {
  FxScanJob j[2];
};

__device__ __forceinline__ const FxScanJob& fxs_job(const FxScanJobs& J, int& b) {
  b = (int)blockIdx.x;
  if (b < J.j[0].nb) return J.j[0];
  b -= J.j[0].nb;
  return J.j[1];
}

extern "C" __global__ __launch_bounds__(FXS_THREADS) void fxScanSums(FxScanJobs J) {
  __shared__ uint lds[FXS_THREADS / 64];
  int b;
  const FxScanJob& S = fxs_job(J, b);
  const uint4 v = fxs_load(S.in, b * FXS_BLOCK + (int)threadIdx.x * FXS_PER, S.n);
  const uint s = fxs_block_sum(v.x + v.y + v.z + v.w, lds);
  if (threadIdx.x == 0) S.sums[b] = s;
}

extern "C" __global__ __launch_bounds__(FXS_THREADS) void fxScanApply(FxScanJobs J) {
  __shared__ uint lds[FXS_THREADS / 64];
  int b;
  const FxScanJob& S = fxs_job(J, b);
  const uint* __restrict__ in = S.in;
  uint* __restrict__ out = S.out;
  const int n = S.n;
  uint pre = 0u;
  for (int j = (int)threadIdx.x; j < b; j += FXS_THREADS) pre += S.sums[j];
  pre = fxs_block_sum(pre, lds);
  const int i = b * FXS_BLOCK + (int)threadIdx.x * FXS_PER;
  const uint4 v = fxs_load(in, i, n);
  const uint t = v.x + v.y + v.z + v.w;
  const uint inc = wave_inclusive_scan(t);
  const int wave = (int)(threadIdx.x >> 6);
  if ((threadIdx.x & 63u) == 63u) lds[wave] = inc;
  __syncthreads();
  uint wpre = 0u;
  for (int w = 0; w < wave; w++) wpre += lds[w];
  const uint e = pre + wpre + inc - t;
  if (i + 3 < n) {
    *reinterpret_cast<uint4*>(out + i) = make_uint4(e, e + v.x, e + v.x + v.y, e + v.x + v.y + v.z);
  } else {
    if (i < n) out[i] = e;
    if (i + 1 < n) out[i + 1] = e + v.x;
    if (i + 2 < n) out[i + 2] = e + v.x + v.y;
  }
}

// ------------------------------------------------------------ counting sort

// x slab of an in-range position inside its cell, from the insert's float
// cell coordinate gx: floor(frac(gx) * FX_SLABS) (gx - (int)gx is exact for
// gx >= 1, and the product by a power of two is exact).
__device__ __forceinline__ int slab_of(const FParams_t& P, float px) {
  const float gx = (px - P.gridMin.x) * P.gridDelta.x;
  const float f = gx - (float)(int)gx;
  return min((int)(f * (float)FX_SLABS), FX_SLABS - 1);
}

// Index sort: AGRID[off[cell] + gndx] = i | slab << FX_MEMBER_BITS, the
// member list of every cell in insertion (atomic) order, each entry carrying
// the member's x slab for the full sort.  Same role as fluids3.2's
// countingSortIndex (fluids3.2/fluids/fluid_system_kern.cu:139-150); reads
// the pre-sort set FPntTmp.  Out-of-range particles fill their scene's tail
// bucket (slab 0).
#define FX_MEMBER_MASK ((1u << FX_MEMBER_BITS) - 1u)
__device__ __forceinline__ void sort_index_impl(const FParams_t& P, const Span S) {
  const int i = S.p0 + S.blk * (int)blockDim.x + (int)threadIdx.x;
  if (i >= S.p1) return;
  const int c = pbuf<const int>(FPntTmp, FGCELL)[i];
  const int cell = (c == FX_GRID_UNDEF) ? S.c0 + P.gridTotal : c;
  const uint* off = pbuf<const uint>(FAccel, AGRIDOFF);
  const uint slot = off[cell] + (uint)pbuf<const int>(FPntTmp, FGNDX)[i];
  const uint slab = c == FX_GRID_UNDEF ? 0u : (uint)slab_of(P, pbuf<const fx_float3>(FPntTmp, FPOS)[i].x);
  pbuf<uint>(FAccel, AGRID)[slot] = (uint)i | (slab << FX_MEMBER_BITS);
}

// Full counting sort FPntTmp -> FPnts (particles.cu:64-95) with a fixed
// in-cell order: a particle's slot in its cell is the number of members that
// precede it by (x slab, pre-sort index) -- the cell split into FX_SLABS
// equal slabs along x (slab_of) -- read from the member list, so the sorted
// order does not depend on atomic timing, and every cell's members lie in
// ascending slabs, which lets the density stage trim each stencil row to the
// particle's x-range.  The tail bucket of out-of-range particles, which can
// hold many, is ranked by the exclusive scan of the insert stage's
// out-of-range flags instead (particle order).  Cells above FX_RANK_CAP
// members keep the atomic order.  The sort moves the state that outlives a
// step (pos, vel, veval, clr, id) and the fresh gcell and gndx; press and
// force are recomputed by the next two stages.
//
// A member entry is (slab << FX_MEMBER_BITS) | index, so "precedes" is one
// unsigned comparison of entries.  Cells of up to FX_RANK_WAVE members are
// ranked per lane that way.  Larger ones are ranked by the wave together, one
// distinct cell at a time: per 64 member entries the wave takes five ballots
// (valid, and the four slab bits) and every lane walks the slab bits from the
// top, keeping the members that agree with its slab so far and counting
// those that fall below it; the members in its own slab with an index below
// the wave's first, and the wave's own lanes below it in the same cell and
// slab, complete the rank.  The same walk with slab = lane gives where each
// slab begins, which lanes 0..FX_SLABS-1 store in ASLABSTART (when the host
// has it) for the density stage; smaller cells get no slab starts.
#ifndef FX_RANK_WAVE
#define FX_RANK_WAVE 16
#endif
#define FX_SLAB_BITS 4
static_assert((1 << FX_SLAB_BITS) == FX_SLABS, "FX_SLABS is a power of two");

__device__ __forceinline__ void sort_full_impl(const FParams_t& P, const Span S) {
  const int lane = (int)(threadIdx.x & 63u);
  const int i = S.p0 + S.blk * (int)blockDim.x + (int)threadIdx.x;
  const bool act = i < S.p1;  // active lanes are a prefix of the wave
  const uint* off = pbuf<const uint>(FAccel, AGRIDOFF);
  const uint* cnt = pbuf<const uint>(FAccel, AGRIDCNT);
  const uint* grid = pbuf<const uint>(FAccel, AGRID);
  const uint* tscan = pbuf<const uint>(FAccel, ATAILSCAN);
  uint* sstart = pbuf<uint>(FAccel, ASLABSTART);
  const int c = act ? __builtin_nontemporal_load(&pbuf<const int>(FPntTmp, FGCELL)[i]) : FX_GRID_UNDEF;
  const int cell = (c == FX_GRID_UNDEF) ? S.c0 + P.gridTotal : c;
  const uint o = act ? off[cell] : 0u;
  const uint n = act ? cnt[cell] : 0u;
  float px = 0.0f, py = 0.0f, pz = 0.0f;
  if (act) {
    const fx_float3 q = ld3nt(pbuf<const fx_float3>(FPntTmp, FPOS), (size_t)i);
    px = q.x;
    py = q.y;
    pz = q.z;
  }
  const int slab = c == FX_GRID_UNDEF ? 0 : slab_of(P, px);
  const uint key = ((uint)slab << FX_MEMBER_BITS) | (uint)i;  // this particle's own member entry
  const bool wide = act && c != FX_GRID_UNDEF && n > FX_RANK_WAVE && n <= FX_RANK_CAP;
  int rank = 0;
  {
    const uint w0 = (uint)(i - lane);
    const int tq = lane & (FX_SLABS - 1);  // the slab whose start this lane counts
    const unsigned long long lower = (1ull << lane) - 1ull;
    unsigned long long pending = __ballot(wide);
    while (pending != 0ull) {  // one distinct dense cell per pass (wave-uniform)
      const int lead = __ffsll((long long)pending) - 1;
      const int cl = __shfl(cell, lead, 64);
      const uint ol = (uint)__shfl((int)o, lead, 64), nl = (uint)__shfl((int)n, lead, 64);
      int below = 0, early = 0, start = 0;
      for (uint k0 = 0; k0 < nl; k0 += 64) {
        const uint k = k0 + (uint)lane;
        const bool valid = k < nl;
        const uint e = valid ? grid[ol + k] : 0u;
        unsigned long long m1 = __ballot(valid), m2 = m1;
        const unsigned long long low = __ballot(valid && (e & FX_MEMBER_MASK) < w0);
#pragma unroll
        for (int b = FX_SLAB_BITS - 1; b >= 0; b--) {
          const unsigned long long bit = __ballot(valid && ((e >> (FX_MEMBER_BITS + b)) & 1u));
          const bool s1 = (slab >> b) & 1, s2 = (tq >> b) & 1;
          below += s1 ? __popcll(m1 & ~bit) : 0;
          m1 = s1 ? (m1 & bit) : (m1 & ~bit);
          start += s2 ? __popcll(m2 & ~bit) : 0;
          m2 = s2 ? (m2 & bit) : (m2 & ~bit);
        }
        early += __popcll(m1 & low);
      }
      const bool mine = wide && cell == cl;
      unsigned long long m3 = __ballot(mine);
#pragma unroll
      for (int b = FX_SLAB_BITS - 1; b >= 0; b--) {
        const unsigned long long bit = __ballot(mine && ((slab >> b) & 1));
        m3 = ((slab >> b) & 1) ? (m3 & bit) : (m3 & ~bit);
      }
      if (mine) rank = below + early + __popcll(m3 & lower);
      if (sstart != nullptr && lane < FX_SLABS) sstart[(size_t)cl * FX_SLABS + lane] = (uint)start;
      pending &= ~__ballot(mine);
    }
  }
  if (!act) return;
  if (wide) {
    // ranked above
  } else if (c == FX_GRID_UNDEF && tscan != nullptr) {
    // tail bucket: the number of the scene's out-of-range particles before i
    rank = (int)(tscan[i] - tscan[S.p0]);
  } else if (n <= FX_RANK_CAP) {
    // member entries below this particle's own; eight in flight per step
    uint k = 0;
    for (; k + 8 <= n; k += 8) {
      uint e[8];
#pragma unroll
      for (int u = 0; u < 8; u++) e[u] = grid[o + k + u];
#pragma unroll
      for (int u = 0; u < 8; u++) rank += e[u] < key ? 1 : 0;
    }
    for (; k < n; k++) rank += grid[o + k] < key ? 1 : 0;
  } else {
    rank = pbuf<const int>(FPntTmp, FGNDX)[i];
  }
  // every load before the first store (the two sets are distinct buffers,
  // which the compiler cannot know); the source set is read once, with
  // non-temporal loads, so that it displaces less of the sorted set that the
  // density stage reads next (sort 0.158 -> 0.146 ms at the 4M beach scene,
  // 0.201 -> 0.192 ms over 200 dam-break steps)
  const bool has_id = FPntTmp.mbuf[FID] != 0 && FPnts.mbuf[FID] != 0;
  const fx_float3 vel = ld3nt(pbuf<const fx_float3>(FPntTmp, FVEL), (size_t)i);
  const fx_float3 vev = ld3nt(pbuf<const fx_float3>(FPntTmp, FVEVAL), (size_t)i);
  const uint clr = __builtin_nontemporal_load(&pbuf<const uint>(FPntTmp, FCLR)[i]);
  const uint id = has_id ? __builtin_nontemporal_load(&pbuf<const uint>(FPntTmp, FID)[i]) : 0u;
  const uint s = o + (uint)rank;
  st3(pbuf<fx_float3>(FPnts, FPOS), s, px, py, pz);
  st3(pbuf<fx_float3>(FPnts, FVEL), s, vel.x, vel.y, vel.z);
  st3(pbuf<fx_float3>(FPnts, FVEVAL), s, vev.x, vev.y, vev.z);
  pbuf<uint>(FPnts, FCLR)[s] = clr;
  if (has_id) pbuf<uint>(FPnts, FID)[s] = id;
  pbuf<int>(FPnts, FGCELL)[s] = c;
  pbuf<int>(FPnts, FGNDX)[s] = rank;
  if (FSortSrc != nullptr) FSortSrc[s] = (uint)i;
}

// The counting sort of a host that runs no countingSortIndex (the reference's
// own, particles.cpp:832-842): every particle goes to off[cell] + gndx, its
// insert slot, so the in-cell order is the insert's atomic order; the eight
// reference buffers move (press and force included, particles.cu:80-89), the
// id when both sets have one, and AGRID[slot] = slot (particles.cu:91).
// Without a tail slot out-of-range particles are not moved (particles.cu:78),
// with one they go to the scene's tail bucket in gndx order.
__device__ __forceinline__ void sort_full_gndx(const FParams_t& P, const Span S, bool tail) {
  const int i = S.p0 + S.blk * (int)blockDim.x + (int)threadIdx.x;
  if (i >= S.p1) return;
  const int c = pbuf<const int>(FPntTmp, FGCELL)[i];
  if (c == FX_GRID_UNDEF && !tail) return;
  const int cell = (c == FX_GRID_UNDEF) ? S.c0 + P.gridTotal : c;
  const int nd = pbuf<const int>(FPntTmp, FGNDX)[i];
  const uint s = pbuf<const uint>(FAccel, AGRIDOFF)[cell] + (uint)nd;
  const bool has_id = FPntTmp.mbuf[FID] != 0 && FPnts.mbuf[FID] != 0;
  const fx_float3 pos = ld3(pbuf<const fx_float3>(FPntTmp, FPOS), (size_t)i);
  const fx_float3 vel = ld3(pbuf<const fx_float3>(FPntTmp, FVEL), (size_t)i);
  const fx_float3 vev = ld3(pbuf<const fx_float3>(FPntTmp, FVEVAL), (size_t)i);
  const fx_float3 frc = ld3(pbuf<const fx_float3>(FPntTmp, FFORCE), (size_t)i);
  const float prs = pbuf<const float>(FPntTmp, FPRESS)[i];
  const uint clr = pbuf<const uint>(FPntTmp, FCLR)[i];
  const uint id = has_id ? pbuf<const uint>(FPntTmp, FID)[i] : 0u;
  st3(pbuf<fx_float3>(FPnts, FPOS), s, pos.x, pos.y, pos.z);
  st3(pbuf<fx_float3>(FPnts, FVEL), s, vel.x, vel.y, vel.z);
  st3(pbuf<fx_float3>(FPnts, FVEVAL), s, vev.x, vev.y, vev.z);
  st3(pbuf<fx_float3>(FPnts, FFORCE), s, frc.x, frc.y, frc.z);
  pbuf<float>(FPnts, FPRESS)[s] = prs;
  pbuf<uint>(FPnts, FCLR)[s] = clr;
  if (has_id) pbuf<uint>(FPnts, FID)[s] = id;
  pbuf<int>(FPnts, FGCELL)[s] = c;
  pbuf<int>(FPnts, FGNDX)[s] = nd;
  uint* grid = pbuf<uint>(FAccel, AGRID);
  if (grid != nullptr) grid[s] = s;
}

extern "C" __global__ void countingSortIndex(int pnum) {
  const int ppt = ppt_of(2);
  const Span S{0, pnum, 0, (int)blockIdx.x};
  if (ppt == 1) {
    sort_index_impl(*cparams(&FParams), S);
    return;
  }
  for (int k = 0; k < ppt; k++) sort_index_impl(*cparams(&FParams), sub_span(S, ppt, k));
}

// The sort of a host that runs the index sort first (libfluidsx).  Kept apart
// from countingSortFull: compiling the insert-slot variant into the same
// kernel slows this one by half (183 vs 114 us per launch at 4M particles).
extern "C" __global__ void countingSortFullRanked(int pnum) {
  const int ppt = ppt_of(2);
  const Span S{0, pnum, 0, xcd_block()};
  if (ppt == 1) {
    sort_full_impl(*cparams(&FParams), S);
    return;
  }
  for (int k = 0; k < ppt; k++) sort_full_impl(*cparams(&FParams), sub_span(S, ppt, k));
}

// The reference's kernel name (particles.cu:64-95): the insert's slots for a
// host without the index sort, the stable ranking for one that runs it
// (FX_CAP_RANKED; countingSortFullRanked is the faster entry for that case).
extern "C" __global__ void countingSortFull(int pnum) {
  const int ppt = ppt_of(2);
  if (!host_ranked()) {
    for (int k = 0; k < ppt; k++)
      sort_full_gndx(*cparams(&FParams), sub_span(Span{0, pnum, 0, (int)blockIdx.x}, ppt, k), host_tail());
    return;
  }
  for (int k = 0; k < ppt; k++) sort_full_impl(*cparams(&FParams), sub_span(Span{0, pnum, 0, xcd_block()}, ppt, k));
}

__device__ __forceinline__ void sort_index_entry(const BatchArgs& B, int b) {
  const int ppt = ppt_of(2);
  const FParams_t* P;
  const Span S = batch_span(B, b, P);
  for (int k = 0; k < ppt; k++) sort_index_impl(*P, sub_span(S, ppt, k));
}

__device__ __forceinline__ void sort_full_entry(const BatchArgs& B, int b) {
  const int ppt = ppt_of(2);
  const FParams_t* P;
  const Span S = batch_span(B, b, P);
  if (ppt == 1) {
    sort_full_impl(*P, S);
    return;
  }
  for (int k = 0; k < ppt; k++) sort_full_impl(*P, sub_span(S, ppt, k));
}

extern "C" __global__ void countingSortIndexBatch(BatchArgs B) { sort_index_entry(B, blockIdx.x); }
extern "C" __global__ void countingSortFullBatch(BatchArgs B) { sort_full_entry(B, xcd_block()); }

extern "C" __global__ void countingSortIndexBatchP(BatchArgs B) {
  batch_items(B, [&](int b) { sort_index_entry(B, b); });
}

extern "C" __global__ void countingSortFullBatchP(BatchArgs B) {
  batch_items(B, [&](int b) { sort_full_entry(B, b); });
}

// ----------------------------------------------------------------- density

// Row decomposition of the stencil: gridAdj[(y*srch + z)*srch + x] =
// (y*res.z + z)*res.x + x (particles.cpp:614-618) is contiguous in x, so the
// srch cells of one (y, z) row are one contiguous run of sorted particles,
// visited in the same order as the reference's per-cell loop.
// Row r of the stencil at base (an in-range particle's): candidates
// [jb, je).  The row ends where the cell after its last one begins -- a cell
// of the same grid row, since an in-range particle's stencil stays one cell
// inside the grid -- so two loads, not three.
__device__ __forceinline__ void row_range(const FParams_t& P, int base, int r, const uint* __restrict__ off,
                                          const uint* __restrict__ cnt, int& jb, int& je) {
  (void)cnt;
  const int srch = P.gridSrch;
  const int c0 = base + P.gridAdj[r * srch];
  jb = (int)off[c0];
  je = (int)off[c0 + srch];
}

__device__ __forceinline__ int stencil_base(const FParams_t& P, int gc) {
  return gc - (P.gridRes.z + 1) * P.gridRes.x - 1;
}

// Tiling.  A block owns the consecutive sorted particles [h0, h1) (in-range
// particles only; out-of-range ones sit in the scene's tail).  Their cells
// form one contiguous index range [cf, cl], so for every stencil row the
// candidates of all the block's particles form one contiguous run of sorted
// particles, from the row start of cf to the row end of cl.  The block stages
// that run in LDS in chunks of TILE_P with coalesced loads; each lane then
// reads the part of the chunk that falls in its own row window.  Chunks are
// visited in increasing order, so every lane still sees its candidates in the
// reference's order.
// One tile entry (a candidate's position) in LDS, 12 bytes: with a 16-byte
// stride entries 16 apart share banks, and the lanes of a wave, whose windows
// start at different entries (their x-range trims differ, and at rest density
// they lie in different cells), queued on them; with a 12-byte stride the
// entries of 21 consecutive candidates fall in distinct banks.  (Separate
// x, y and z arrays measured the same as this.)
struct TileP {
  float x, y, z;
};
#ifndef TILE_P
#define TILE_P 1024  // candidates per tile, 12 KB
#endif
#define FX_TILE_FLOATS (3 * TILE_P)
__device__ __forceinline__ void tile_put(float* t, int k, const fx_float3& q) {
  t[3 * k] = q.x;
  t[3 * k + 1] = q.y;
  t[3 * k + 2] = q.z;
}
__device__ __forceinline__ TileP tile_get(const float* t, int k) { return TileP{t[3 * k], t[3 * k + 1], t[3 * k + 2]}; }

struct BlockSpan {
  int h1;      // end of the block's in-range particles
  int cf, cl;  // first and last cell of the block
};

__device__ __forceinline__ bool block_span(const FParams_t& P, const Span S, int h0,
                                           const int* __restrict__ gcell, const uint* __restrict__ off,
                                           const uint* __restrict__ cnt, bool tail, BlockSpan& s) {
  // end of the scene's in-range particles: the start of its tail bucket, or
  // the end of its last cell when the table has no tail slot
  const int last = S.c0 + P.gridTotal - 1;
  const int nvalid = tail ? (int)off[last + 1] : (int)(off[last] + cnt[last]);
  if (h0 >= nvalid || h0 >= S.p1) return false;
  s.h1 = min(min(h0 + (int)blockDim.x, nvalid), S.p1);
  s.cf = gcell[h0];
  s.cl = gcell[s.h1 - 1];
  return true;
}

// Number of bitmask words a particle of cell gc uses: one word-aligned run of
// ceil(len/32) words per stencil row.
__device__ __forceinline__ int nbr_words(const FParams_t& P, int gc, const uint* __restrict__ off,
                                         const uint* __restrict__ cnt) {
  const int base = stencil_base(P, gc);
  int w = 0;
  for (int r = 0; r < P.gridSrch * P.gridSrch; r++) {
    int jb, je;
    row_range(P, base, r, off, cnt, jb, je);
    w += (je - jb + 31) >> 5;
  }
  return w;
}

// Density (particles.cu:120-157): rho_i = m * poly6 * sum_{0<d2<rd2}
// ((rd2 - d2) * d2s)^3 with d2 in world units, stored in FPRESS; an isolated
// particle gets 1.0.
//
// When the host provides the ANBRBITS and AFREC buffers the kernel also
// records, for the force stage, one bit per stencil candidate (set for the
// in-radius pairs; row r of particle i occupies words W_r .. W_r+ceil(len/32)-1
// at [w * N + i], N the size of the whole particle buffer) and the particle's
// (pos, rho), (veval, 0) record.  A particle whose candidates need more than
// the cap's words (FNbrWords) records only its first ones, which the force
// stage does not read: it finds the particle over the cap from its own row
// lengths and tests its candidates in place.
//
// Row trimming (when the sort ordered every cell's members by x slab and
// wrote where each slab begins, `trim`).  In grid units (the insert's
// gx = (x - gridMin.x) * gridDelta.x, cells of width 1) a member of slab k of
// a cell at x index c has gx in [c + k/16, c + (k+1)/16).  A particle at gp
// is out of radius of a whole slab of its row's first cell (x index cx - 1)
// when gp - (cx - 1) - (k+1)/16 > Rg, and of a whole slab of its row's last
// cell (x index cx + srch - 2) when (cx + srch - 2) + k/16 - gp > Rg, with
// Rg = rlim * gridDelta.x plus a margin of 2^-7 cells and gridRes.x * 2^-20
// that covers the rounding of gx (about 2^-22 of it) on both sides: every
// member of such a slab has fl(|px - x|) > rlim, so fl(dx * dx) > rd2 and it
// fails the radius test whatever its y and z.  Those slabs are a prefix of
// the first cell and a suffix of the last (the members lie in ascending
// slabs), located through the slab starts, and the lane walks only [ts, te)
// of its row.  The skipped candidates' bits are zero: the bitmask keeps its
// layout (bit s of the row for candidate s), written with the lane's phase
// ts % 32.
#define FX_TRIM_MARGIN 0.0078125f

// Slabs [0, lo) of the row's first cell and [hi, FX_SLABS) of its last cell
// lie beyond the radius of a particle at grid x coordinate gp in cell cx.
__device__ __forceinline__ void trim_slabs(float gp, int cx, int srch, float Rg, int& lo, int& hi) {
  const float fx = (float)(cx - 1), lx = (float)(cx + srch - 2);
  lo = min(max((int)ceilf((float)FX_SLABS * (gp - fx - Rg) - 1.0f), 0), FX_SLABS);
  hi = min(max((int)floorf((float)FX_SLABS * (gp - lx + Rg)) + 1, 0), FX_SLABS);
}

// Members of cell c (n of them) in slabs below k.
__device__ __forceinline__ int slab_start(const uint* __restrict__ sstart, int c, int n, int k) {
  return (n == 0 || k == 0) ? 0 : (k >= FX_SLABS ? n : (int)sstart[(size_t)c * FX_SLABS + k]);
}

__device__ __forceinline__ void pressure_impl(const FParams_t& P, const Span S, int N, float* tile, bool tail,
                                              bool trim) {
  const int t = threadIdx.x;
  const int h0 = S.p0 + S.blk * (int)blockDim.x;
  const int i = h0 + t;
  const int* __restrict__ gcell = pbuf<const int>(FPnts, FGCELL);
  const fx_float3* __restrict__ pos = pbuf<const fx_float3>(FPnts, FPOS);
  const uint* __restrict__ off = pbuf<const uint>(FAccel, AGRIDOFF);
  const uint* __restrict__ cnt = pbuf<const uint>(FAccel, AGRIDCNT);
  uint* __restrict__ bits = pbuf<uint>(FAccel, ANBRBITS);
  const bool wbits = bits != nullptr && FAccel.mbuf[AFREC] != 0;
  BlockSpan B;
  if (!block_span(P, S, h0, gcell, off, cnt, tail, B)) return;  // uniform per block
  const bool act = i < B.h1;
  const int gc = act ? gcell[i] : B.cf;
  fx_float3 p = fx_float3{0.0f, 0.0f, 0.0f};
  if (act) p = pos[i];
  const int shift = (P.gridRes.z + 1) * P.gridRes.x + 1;
  const int srch = P.gridSrch;
  const float rd2 = P.rd2, d2 = P.d2;
  const size_t NN = (size_t)N;
  // The bitmask's words are written up to the cap: a particle that needs
  // more is found so by the force stage from its own row lengths, which then
  // tests its candidates in place and reads none of them.
  const bool rec = wbits && act;
  const int cap = nbr_cap();
  auto put = [&](int w, uint v) {
    if (rec && w < cap) bits[(size_t)w * NN + i] = v;
  };
  // x-distance beyond which fl(dx * dx) > rd2: dx > rlim > sqrt(rd2) (1 + 2^-21)
  const float rlim = __fsqrt_rn(rd2) * (1.0f + 0x1p-20f);
  const float Rg = rlim * P.gridDelta.x + FX_TRIM_MARGIN + (float)P.gridRes.x * 0x1p-20f;
  const float gp = (p.x - P.gridMin.x) * P.gridDelta.x;  // the insert's cell coordinate
  const uint* __restrict__ sstart = pbuf<const uint>(FAccel, ASLABSTART);
  const bool trim_lane = trim && sstart != nullptr && act;
  int slo = 0, shi = FX_SLABS;
  if (trim_lane) trim_slabs(gp, (int)gp, srch, Rg, slo, shi);
  float sum = 0.0f;
  int W = 0;  // first word of the current row
  for (int r = 0; r < srch * srch; r++) {
    const int adj = P.gridAdj[r * srch] - shift;
    // a row ends where the cell after its last one begins (an in-range
    // particle's row is followed by a cell of the same grid row)
    const int lo = (int)off[B.cf + adj];
    const int hi = (int)off[B.cl + adj + srch];
    int jb = 0, je = 0;
    if (act) {
      jb = (int)off[gc + adj];
      je = (int)off[gc + adj + srch];
    }
    // the lane walks candidates [ts, te) of its row (relative to jb); a row
    // of at most FX_RANK_WAVE candidates has no cell with slab starts
    int ts = 0, te = je - jb;
    if (trim_lane && je - jb > FX_RANK_WAVE) {
      const int c0 = gc + adj, c2 = gc + adj + srch - 1;
      const int n0 = (int)cnt[c0], n2 = (int)cnt[c2];
      // slab starts exist for the cells the sort ranked by wave
      if (n0 > FX_RANK_WAVE && n0 <= FX_RANK_CAP) ts = slab_start(sstart, c0, n0, slo);
      if (n2 > FX_RANK_WAVE && n2 <= FX_RANK_CAP) te = max(ts, te - n2 + slab_start(sstart, c2, n2, shi));
    }
    if (hi - lo <= TILE_P) {
      // The row's run fits one tile: the lane's candidate s (0 <= s < len)
      // is tile entry a + s and bit s % 32 of the row's word s / 32, the
      // same position in every lane.  The wave walks s in lockstep, 32
      // candidates per word; candidates every lane has are unchecked, the
      // rest are masked by s < len.  Bits are shifted in (first candidate
      // highest) and the word is bit-reversed when stored.
      for (int k = t; k < hi - lo; k += blockDim.x) {
        const fx_float3 q = pos[lo + k];
        tile_put(tile, k, q);
      }
      __syncthreads();
      const int a = act ? jb - lo : 0, len = act ? je - jb : 0;
      const int L = te - ts, a0 = a + ts, ph = ts & 31, w0 = ts >> 5;
      const int lmax = wave_max(L), lmin = wave_min(act ? L : INT_MAX);
      const int last = hi - lo - 1;  // last entry of the run
      const int nw = (len + 31) >> 5;
      uint acc = 0u;
      auto cand = [&](int s, auto checked) {
        int k = a0 + s;
        if (decltype(checked)::value) k = min(k, last);
        const TileP q = tile_get(tile, k);
        const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        const float dsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        if (decltype(checked)::value) {
          const bool pair = dsq < rd2 && dsq > 0.0f && s < L;
          const float tt = pair ? (rd2 - dsq) * d2 : 0.0f;
          sum = fmaf(tt * tt, tt, sum);  // + 0 when not a pair: sum unchanged
          acc = acc + acc + (pair ? 1u : 0u);
        } else {
          // The same three steps with the wave's pair mask m used as it is:
          // the term selected by m (v_cndmask) and the bit shifted in as the
          // carry of acc + acc (v_addc), one instruction each where the
          // compiler spends two on the bit.  One ballot per compare, so that
          // m is the compares' own mask.
          const unsigned long long m = __builtin_amdgcn_ballot_w64(dsq < rd2) & __builtin_amdgcn_ballot_w64(dsq > 0.0f);
          const float t0 = (rd2 - dsq) * d2;
          float tt;
          __asm__("v_cndmask_b32_e64 %0, 0, %1, %2" : "=v"(tt) : "v"(t0), "s"(m));
          sum = fmaf(tt * tt, tt, sum);  // + 0 when not a pair: sum unchanged
          unsigned long long carry;
          uint acc2;
          __asm__("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(acc2), "=s"(carry) : "v"(acc), "s"(m));
          acc = acc2;
        }
      };
      // the block of candidates ts + s0 .. ts + s0 + 31 straddles words
      // w0 + s0/32 (bits ph..31) and the next (bits 0..ph-1): the second
      // part is carried into the next block's word
      uint carry = 0u;
      int wb = w0;
      for (int s0 = 0; s0 < lmax; s0 += 32) {
        const int n = min(32, lmax - s0);
        const int nu = max(0, min(n, lmin - s0));
        int u = 0;
        // unrolled by hand: the asm in cand keeps the compiler from unrolling
        for (; u + 4 <= nu; u += 4) {
          cand(s0 + u, std::false_type());
          cand(s0 + u + 1, std::false_type());
          cand(s0 + u + 2, std::false_type());
          cand(s0 + u + 3, std::false_type());
        }
        for (; u < nu; u++) cand(s0 + u, std::false_type());
        for (; u < n; u++) cand(s0 + u, std::true_type());
        const uint chunk = __builtin_bitreverse32(acc << (32 - n));
        if (wb < nw) put(W + wb, carry | (chunk << ph));
        carry = ph ? chunk >> (32 - ph) : 0u;
        wb++;
        acc = 0u;
      }
      if (rec) {  // the words the walk did not reach hold no pair
        for (int w = 0; w < w0; w++) put(W + w, 0u);
        if (wb < nw) put(W + wb, carry);
        for (int w = wb + 1; w < nw; w++) put(W + w, 0u);
      }
      __syncthreads();
      W += (je - jb + 31) >> 5;
      continue;
    }
    // bits of the current word and the next one's mask, from the lane's phase
    uint acc = 0, bit = 1u << (ts & 31);
    int wcur = W + (ts >> 5);  // current word
    if (rec)
      for (int w = W; w < wcur; w++) put(w, 0u);  // skipped candidates
    for (int c0 = lo; c0 < hi; c0 += TILE_P) {
      const int n = min(TILE_P, hi - c0);
      for (int k = t; k < n; k += blockDim.x) {
        const fx_float3 q = pos[c0 + k];
        tile_put(tile, k, q);
      }
      __syncthreads();
      const int a = max(jb + ts, c0) - c0, b = min(jb + te, c0 + n) - c0;
      for (int k = a; k < b; k++) {
        const TileP q = tile_get(tile, k);
        const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        const float dsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        const bool pair = dsq < rd2 && dsq > 0.0f;
        if (pair) {
          const float tt = (rd2 - dsq) * d2;
          sum = fmaf(tt * tt, tt, sum);
        }
        if (rec) {
          acc |= pair ? bit : 0u;
          if (bit == 0x80000000u) {  // word complete
            put(wcur, acc);
            acc = 0;
            wcur++;
          }
          bit = __builtin_amdgcn_alignbit(bit, bit, 31);  // rotate left by one
        }
      }
      __syncthreads();
    }
    if (rec && bit != 1u) put(wcur++, acc);  // the walk's partial last word
    if (rec)
      for (int w = wcur; w < W + ((je - jb + 31) >> 5); w++) put(w, 0u);
    W += (je - jb + 31) >> 5;
  }
  if (act) {
    sum = sum * P.pmass * P.poly6kern;
    if (sum == 0.0f) sum = 1.0f;
    pbuf<float>(FPnts, FPRESS)[i] = sum;
    if (wbits) {
      // read once here: non-temporal (density 0.305 -> 0.298 ms at the 4M beach scene)
      const fx_float3 v = ld3nt(pbuf<const fx_float3>(FPnts, FVEVAL), (size_t)i);
      float4* fr = pbuf<float4>(FAccel, AFREC);
      fr[2 * (size_t)i] = make_float4(p.x, p.y, p.z, sum);
      fr[2 * (size_t)i + 1] = make_float4(v.x, v.y, v.z, 0.0f);
    }
  }
}

extern "C" __global__ __launch_bounds__(1024) void computePressure(int pnum) {
  __shared__ __align__(16) float tile[FX_TILE_FLOATS];
  pressure_impl(*cparams(&FParams), Span{0, pnum, 0, xcd_block()}, pnum, tile, host_tail(), host_ranked());
}

extern "C" __global__ __launch_bounds__(1024) void computePressureBatch(BatchArgs B, int ntotal) {
  __shared__ __align__(16) float tile[FX_TILE_FLOATS];
  const FParams_t* P;
  const Span S = batch_span(B, xcd_block(), P);
  pressure_impl(*P, S, ntotal, tile, true, true);
}

extern "C" __global__ __launch_bounds__(1024) void computePressureBatchP(BatchArgs B, int ntotal) {
  __shared__ __align__(16) float tile[FX_TILE_FLOATS];
  batch_items(B, [&](int b) {
    const FParams_t* P;
    const Span S = batch_span(B, b, P);
    pressure_impl(*P, S, ntotal, tile, true, true);
    __syncthreads();  // the next entry's tile loads wait for this entry's last reads
  });
}

// ------------------------------------------------------------------- force

// Force (particles.cu:188-225): for 0 < d2 < rd2, r = sqrt(d2 * d2s),
// c = h - r, p = (rho - rho0) * k,
// pterm = (ss * -0.5) * c * spiky * (pi + pj) / r,
// F += (pterm * dist + vterm * (veval_j - veval_i)) * c / (rho_i rho_j).
//
// With the density stage's bitmask the kernel never re-tests candidates.
// Two walks over the bitmask: the word queue further below (blocks of up to
// 256 threads, 2- and 3-cell stencil rows), and the pair ring here (any
// block size and stencil).  The ring: each lane turns the set bits of its
// words into neighbour indices in a per-lane ring in LDS; the wave then
// evaluates queued pairs in batches: a batch takes from every lane as many
// pairs as the shortest active queue holds (more only when a queue must make
// room).  Each lane adds its own pairs in candidate order.  Waves are
// independent: no block-wide barrier.
#define FQ_SLOTS 1024  // ring slots per block (16 per lane at 64 threads)

struct ForceI {
  float px, py, pz, vx, vy, vz;
  float di, pi;
  float fx, fy, fz;
};

// One pair's term, arranged around a single correctly rounded reciprocal:
// with D = rho_i rho_j and inv = 1 / (r D),
//   F += A dist + B (veval_j - veval_i),
//   A = ((ss * -0.5) * spiky) c^2 (p_i + p_j) inv,  B = vterm c r inv,
// which is the reference's (pterm dist + vterm dv) c / D with
// pterm = ss (-0.5) c spiky (p_i + p_j) / r.  The term is split from its
// accumulation so that several pairs' terms can be evaluated before they
// are added in order.
struct PairTerm {
  float A, B, dx, dy, dz, ux, uy, uz;
};

__device__ __forceinline__ PairTerm pair_term(const FParams_t& P, const ForceI& I, const float4 a,
                                              const float4 v) {
  PairTerm q;
  q.dx = I.px - a.x;
  q.dy = I.py - a.y;
  q.dz = I.pz - a.z;
  const float dsq = fmaf(q.dz, q.dz, fmaf(q.dy, q.dy, q.dx * q.dx));
  const float rr = sqrtf(dsq * P.d2);
  const float c = P.psmoothradius - rr;
  const float dj = a.w;
  const float pj = (dj - P.prest_dens) * P.pintstiff;
  const float inv = 1.0f / (rr * (I.di * dj));
  const float kp = (P.sim_scale * -0.5f) * P.spikykern;
  q.A = (kp * (c * c)) * (I.pi + pj) * inv;
  q.B = (P.vterm * c) * rr * inv;
  q.ux = v.x - I.vx;
  q.uy = v.y - I.vy;
  q.uz = v.z - I.vz;
  return q;
}

__device__ __forceinline__ void pair_add(ForceI& I, const PairTerm& q) {
  I.fx = fmaf(q.A, q.dx, fmaf(q.B, q.ux, I.fx));
  I.fy = fmaf(q.A, q.dy, fmaf(q.B, q.uy, I.fy));
  I.fz = fmaf(q.A, q.dz, fmaf(q.B, q.uz, I.fz));
}

// pair_add when ok, else nothing (a select, no branch)
__device__ __forceinline__ void pair_add_if(ForceI& I, const PairTerm& q, bool ok) {
  const float fx = fmaf(q.A, q.dx, fmaf(q.B, q.ux, I.fx));
  const float fy = fmaf(q.A, q.dy, fmaf(q.B, q.uy, I.fy));
  const float fz = fmaf(q.A, q.dz, fmaf(q.B, q.uz, I.fz));
  I.fx = ok ? fx : I.fx;
  I.fy = ok ? fy : I.fy;
  I.fz = ok ? fz : I.fz;
}

__device__ __forceinline__ void force_pair(const FParams_t& P, ForceI& I, const float4 a, const float4 v) {
  pair_add(I, pair_term(P, I, a, v));
}

__device__ __forceinline__ bool is_pair(const FParams_t& P, const ForceI& I, const float4 a) {
  const float dx = I.px - a.x, dy = I.py - a.y, dz = I.pz - a.z;
  const float dsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  return dsq < P.rd2 && dsq > 0.0f;
}

__device__ __forceinline__ void force_impl(const FParams_t& P, const Span S, int N, uint* fq) {
  const int t = threadIdx.x;
  const int i = S.p0 + S.blk * (int)blockDim.x + t;
  const uint* __restrict__ off = pbuf<const uint>(FAccel, AGRIDOFF);
  const uint* __restrict__ cnt = pbuf<const uint>(FAccel, AGRIDCNT);
  const uint* __restrict__ bits = pbuf<const uint>(FAccel, ANBRBITS);
  const float4* __restrict__ fr = pbuf<const float4>(FAccel, AFREC);
  const int srch = P.gridSrch;
  int gc = FX_GRID_UNDEF;
  if (i < S.p1) gc = pbuf<const int>(FPnts, FGCELL)[i];
  const bool act = gc != FX_GRID_UNDEF;
  const int base = stencil_base(P, gc);
  ForceI I;
  I.px = I.py = I.pz = I.vx = I.vy = I.vz = 0.0f;
  I.di = 1.0f;
  I.pi = 0.0f;
  I.fx = I.fy = I.fz = 0.0f;

  if (bits == nullptr || fr == nullptr) {
    // no density-stage records (a host without ANBRBITS/AFREC): test every
    // candidate and evaluate pairs in place
    if (act) {
      const fx_float3* __restrict__ pos = pbuf<const fx_float3>(FPnts, FPOS);
      const fx_float3* __restrict__ vev = pbuf<const fx_float3>(FPnts, FVEVAL);
      const float* __restrict__ rho = pbuf<const float>(FPnts, FPRESS);
      const fx_float3 p = pos[i], v = vev[i];
      I.px = p.x; I.py = p.y; I.pz = p.z;
      I.vx = v.x; I.vy = v.y; I.vz = v.z;
      I.di = rho[i];
      I.pi = (I.di - P.prest_dens) * P.pintstiff;
      for (int r = 0; r < srch * srch; r++) {
        int jb, je;
        row_range(P, base, r, off, cnt, jb, je);
        for (int j = jb; j < je; j++) {
          const fx_float3 q = pos[j], u = vev[j];
          const float4 a = make_float4(q.x, q.y, q.z, rho[j]);
          if (is_pair(P, I, a)) force_pair(P, I, a, make_float4(u.x, u.y, u.z, 0.0f));
        }
      }
      pbuf<fx_float3>(FPnts, FFORCE)[i] = fx_float3{I.fx, I.fy, I.fz};
    }
    return;
  }

  if (act) {
    const float4 a = fr[2 * (size_t)i], v = fr[2 * (size_t)i + 1];
    I.px = a.x; I.py = a.y; I.pz = a.z;
    I.vx = v.x; I.vy = v.y; I.vz = v.z;
    I.di = a.w;
    I.pi = (I.di - P.prest_dens) * P.pintstiff;
  }
  // a particle without recorded bits (too many candidates) tests its own
  const bool ovf = act && nbr_words(P, gc, off, cnt) > nbr_cap();
  if (ovf) {
    for (int r = 0; r < srch * srch; r++) {
      int jb, je;
      row_range(P, base, r, off, cnt, jb, je);
      for (int j = jb; j < je; j++) {
        const float4 a = fr[2 * (size_t)j];
        if (is_pair(P, I, a)) force_pair(P, I, a, fr[2 * (size_t)j + 1]);
      }
    }
  }
  const bool use = act && !ovf;
  const size_t NN = (size_t)N;
  const int fqn = FQ_SLOTS / (int)blockDim.x;  // ring entries per lane
  if (fqn < 16) {
    // blocks too large for a ring that holds a half-word per lane: evaluate
    // each lane's pairs in place
    int W = 0;
    for (int r = 0; r < srch * srch; r++) {
      int jb = 0, je = 0;
      if (use) row_range(P, base, r, off, cnt, jb, je);
      const int nw = (je - jb + 31) >> 5;
      for (int w = 0; w < nw; w++) {
        uint m = bits[(size_t)(W + w) * NN + i];
        while (m != 0u) {
          const uint j = (uint)(jb + (w << 5) + __builtin_ctz(m));
          m &= m - 1u;
          force_pair(P, I, fr[2 * (size_t)j], fr[2 * (size_t)j + 1]);
        }
      }
      W += nw;
    }
    if (act) pbuf<fx_float3>(FPnts, FFORCE)[i] = fx_float3{I.fx, I.fy, I.fz};
    return;
  }
  int head = 0, qn = 0;  // ring holds entries head .. head+qn-1 (mod fqn)
  int W = 0;
  // evaluates the ring's first f entries of every lane (fewer if it holds
  // fewer); a lane whose ring is empty evaluates record 0 and discards the
  // result.
  auto drain = [&](int f) {
#pragma unroll 2
    for (int e = 0; e < f; e++) {
      const bool ok = e < qn;
      const uint j = ok ? fq[head * (int)blockDim.x + t] : 0u;
      const float4 a = fr[2 * (size_t)j], v = fr[2 * (size_t)j + 1];
      ForceI J = I;
      force_pair(P, J, a, v);
      if (ok) {
        I.fx = J.fx;
        I.fy = J.fy;
        I.fz = J.fz;
        head = (head + 1 == fqn) ? 0 : head + 1;
      }
    }
    qn = max(qn - f, 0);
  };
  for (int r = 0; r < srch * srch; r++) {
    int jb = 0, je = 0;
    if (use) row_range(P, base, r, off, cnt, jb, je);
    const int nw = (je - jb + 31) >> 5;
    const int maxw = min(wave_max(nw), FX_NBR_WORDS);
    for (int w = 0; w < maxw; w++) {
      uint word = 0;
      if (w < nw) word = bits[(size_t)(W + w) * NN + i];
      for (int h = 0; h < 32; h += 16) {  // a word in two halves of 16 candidates
        uint m = (word >> h) & 0xffffu;
        const int pc = __popc(m);
        // make room: every lane must fit this half-word's pairs
        if (__any(qn + pc > fqn)) {
          const int need = wave_max(qn + pc - fqn);
          drain(min(max(need, wave_min(use ? qn : 0x7fffffff)), fqn));
        }
        int tail = head + qn;
        if (tail >= fqn) tail -= fqn;
        while (m != 0u) {
          const int b = __builtin_ctz(m);
          m &= m - 1u;
          fq[tail * (int)blockDim.x + t] = (uint)(jb + (w << 5) + h + b);
          tail = (tail + 1 == fqn) ? 0 : tail + 1;
          qn++;
        }
      }
    }
    W += nw;
  }
  drain(min(wave_max(qn), fqn));
  if (act) pbuf<fx_float3>(FPnts, FFORCE)[i] = fx_float3{I.fx, I.fy, I.fz};
}

#ifndef FW_AHEAD
#define FW_AHEAD 4  // bitmask words loaded ahead of their use
#endif

// Word-queue force.  Each lane reads its bitmask words as one stream (row
// r's words follow row r-1's) and queues every non-zero word whole, with the
// record index of its bit 0: two words in registers, up to K more in LDS.
// The wave then evaluates pairs in batches: a batch takes the same number of
// pairs from every lane (as many as the shortest queue holds, more only when
// a queue must make room), each lane's pairs in candidate order, FW_STEP
// pairs' terms evaluated at a time.  Queuing words instead of pairs costs one LDS
// write per word, and a deep queue of words keeps nearly every lane busy in
// the batches, so nearly every pair gather carries a pair.
#ifndef FW_STEP
#define FW_STEP 3  // pair terms evaluated per drain step
#endif
#ifndef FW_QBYTES
#define FW_QBYTES 16384  // LDS queue per block: (word, base) pairs, 4 words per lane at 512 threads
#endif

template <int R, int S>
__device__ __forceinline__ void force_walk_q(const FParams_t& P, ForceI& I, const uint* __restrict__ bits,
                                             size_t NN, int i, int nwords, int (&E_)[R], int (&G)[R], uint* Qw,
                                             uint* Qb, int K, const float4* srcA, const float4* srcV, bool use) {
  const int t = threadIdx.x, nt = blockDim.x;
  uint cur = 0u, nxt = 0u;  // current word (bits not yet taken) and the next one
  int curb = 0, nxtb = 0;   // their bases
  int qh = 0, qc = 0;       // LDS queue head and length
  int qn = 0;               // pairs held
  // the lane's next pair, in candidate order; false when it holds none
  auto take = [&](uint& j) -> bool {
    const bool ok = cur != 0u;
    j = ok ? (uint)(curb + __builtin_ctz(cur)) : 0u;
    cur &= cur - 1u;
    if (ok && cur == 0u) {
      cur = nxt;
      curb = nxtb;
      nxt = 0u;
      if (qc > 0) {
        nxt = Qw[qh * nt + t];
        nxtb = (int)Qb[qh * nt + t];
        qh = (qh + 1 == K) ? 0 : qh + 1;
        qc--;
      }
    }
    return ok;
  };
  // f pairs from every lane (a lane that runs out evaluates record 0 and
  // discards it), FW_STEP pair terms evaluated at a time and added in order
  auto drain = [&](int f) {
    int e = 0;
    for (; e + FW_STEP <= f; e += FW_STEP) {
      uint j[FW_STEP];
      bool ok[FW_STEP];
      PairTerm q[FW_STEP];
#pragma unroll
      for (int u = 0; u < FW_STEP; u++) ok[u] = take(j[u]);
#pragma unroll
      for (int u = 0; u < FW_STEP; u++) q[u] = pair_term(P, I, srcA[S * (size_t)j[u]], srcV[S * (size_t)j[u]]);
#pragma unroll
      for (int u = 0; u < FW_STEP; u++) pair_add_if(I, q[u], ok[u]);
    }
    for (; e < f; e++) {
      uint j;
      const bool ok = take(j);
      pair_add_if(I, pair_term(P, I, srcA[S * (size_t)j], srcV[S * (size_t)j]), ok);
    }
    qn = max(qn - f, 0);
  };
  const int maxf = wave_max(nwords);
  // The bitmask is read once, as a stream: non-temporal loads keep it from
  // displacing the neighbour records in L1 and L2 (force 3.81 -> 3.79 ms per
  // launch over 200 dam-break steps, 0.470 -> 0.451 ms at the 4M beach scene).
  // Non-temporal stores in the density stage measured slower (3.85 ms).
  uint nx4[FW_AHEAD];
#pragma unroll
  for (int k = 0; k < FW_AHEAD; k++) nx4[k] = k < nwords ? __builtin_nontemporal_load(&bits[(size_t)k * NN + i]) : 0u;
  for (int f0 = 0; f0 < maxf; f0 += FW_AHEAD) {
    uint cw[FW_AHEAD];
#pragma unroll
    for (int k = 0; k < FW_AHEAD; k++) {
      cw[k] = nx4[k];
      const int fn = f0 + FW_AHEAD + k;
      nx4[k] = fn < nwords ? __builtin_nontemporal_load(&bits[(size_t)fn * NN + i]) : 0u;
    }
#pragma unroll
    for (int k = 0; k < FW_AHEAD; k++) {
      const int f = f0 + k;
      if (f >= maxf) break;
      const uint word = cw[k];
      if (word != 0u) {
        while (f >= E_[0]) {  // move to the row holding word f
#pragma unroll
          for (int r = 0; r < R - 1; r++) {
            E_[r] = E_[r + 1];
            G[r] = G[r + 1];
          }
          E_[R - 1] = INT_MAX;
        }
      }
      const int held = (cur != 0u ? 1 : 0) + (nxt != 0u ? 1 : 0) + qc;
      const bool full = word != 0u && held == K + 2;
      if (__any(full)) {  // free a word slot: the full lanes finish their current word
        const int need = wave_max(full ? __popc(cur) : 0);
        drain(max(need, wave_min(use ? qn : 0x7fffffff)));
      }
      if (word != 0u) {
        const int base = G[0] + (f << 5);
        if (cur == 0u) {
          cur = word;
          curb = base;
        } else if (nxt == 0u) {
          nxt = word;
          nxtb = base;
        } else {
          int tl = qh + qc;
          if (tl >= K) tl -= K;
          Qw[tl * nt + t] = word;
          Qb[tl * nt + t] = (uint)base;
          qc++;
        }
        qn += __popc(word);
      }
    }
  }
  drain(wave_max(qn));
}

template <int SR>
__device__ __forceinline__ void force_q_impl(const FParams_t& P, const Span S, int N, uint* lq) {
  constexpr int R = SR * SR;
  const int t = threadIdx.x, nt = blockDim.x;
  const int i = S.p0 + S.blk * nt + t;
  const uint* __restrict__ off = pbuf<const uint>(FAccel, AGRIDOFF);
  const uint* __restrict__ cnt = pbuf<const uint>(FAccel, AGRIDCNT);
  const uint* __restrict__ bits = pbuf<const uint>(FAccel, ANBRBITS);
  const float4* __restrict__ fr = pbuf<const float4>(FAccel, AFREC);
  int gc = FX_GRID_UNDEF;
  if (i < S.p1) gc = pbuf<const int>(FPnts, FGCELL)[i];
  const bool act = gc != FX_GRID_UNDEF;
  const int shift = (P.gridRes.z + 1) * P.gridRes.x + 1;
  ForceI I;
  I.px = I.py = I.pz = I.vx = I.vy = I.vz = 0.0f;
  I.di = 1.0f;
  I.pi = 0.0f;
  I.fx = I.fy = I.fz = 0.0f;
  // the lane's row windows: row r covers words [W_r, W_{r+1}) of its bitmask;
  // E_[r] = W_{r+1}, and the record of word f, bit b is G[r] + 32 f + b
  int E_[R], G[R];
  int nwords = 0;
  if (act) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int c0 = gc + P.gridAdj[r * SR] - shift;
      const int jb = (int)off[c0], je = (int)off[c0 + SR];
      G[r] = jb - (nwords << 5);
      nwords += (je - jb + 31) >> 5;
      E_[r] = nwords;
    }
    const float4 a = fr[2 * (size_t)i], v = fr[2 * (size_t)i + 1];
    I.px = a.x; I.py = a.y; I.pz = a.z;
    I.vx = v.x; I.vy = v.y; I.vz = v.z;
    I.di = a.w;
    I.pi = (I.di - P.prest_dens) * P.pintstiff;
  } else {
#pragma unroll
    for (int r = 0; r < R; r++) E_[r] = G[r] = 0;
  }
  const bool ovf = act && nwords > nbr_cap();
  if (ovf) {  // no recorded bits: test every candidate
    const int base = stencil_base(P, gc);
    for (int r = 0; r < R; r++) {
      int jb, je;
      row_range(P, base, r, off, cnt, jb, je);
      for (int j = jb; j < je; j++) {
        const float4 a = fr[2 * (size_t)j];
        if (is_pair(P, I, a)) force_pair(P, I, a, fr[2 * (size_t)j + 1]);
      }
    }
  }
  const bool use = act && !ovf;
  const int K = min(FW_QBYTES / (8 * nt), 16);
  force_walk_q<R, 2>(P, I, bits, (size_t)N, i, use ? nwords : 0, E_, G, lq, lq + K * nt, K, fr, fr + 1, use);
  if (act) pbuf<fx_float3>(FPnts, FFORCE)[i] = fx_float3{I.fx, I.fy, I.fz};
}

// Dispatch: the word queue needs the density stage's records and at least
// 4 LDS words per lane (blocks of up to 512 threads, the reference's launch
// geometry); otherwise the ring or candidate-testing pass.
__device__ __forceinline__ void force_q(const FParams_t& P, const Span S, int N, uint* lq) {
  if (FAccel.mbuf[ANBRBITS] == 0 || FAccel.mbuf[AFREC] == 0 || FW_QBYTES / (8 * (int)blockDim.x) < 4 ||
      P.gridSrch > 3) {
    force_impl(P, S, N, lq);
    return;
  }
  if (P.gridSrch == 3)
    force_q_impl<3>(P, S, N, lq);
  else
    force_q_impl<2>(P, S, N, lq);
}

extern "C" __global__ __launch_bounds__(1024) void computeForce(int pnum) {
  __shared__ uint lq[FW_QBYTES / 4];
  force_q(*cparams(&FParams), Span{0, pnum, 0, xcd_block()}, pnum, lq);
}

extern "C" __global__ __launch_bounds__(1024) void computeForceBatch(BatchArgs B, int ntotal) {
  __shared__ uint lq[FW_QBYTES / 4];
  const FParams_t* P;
  const Span S = batch_span(B, xcd_block(), P);
  force_q(*P, S, ntotal, lq);
}

// The persistent loop costs the kernel registers (101 VGPRs, 4 waves per
// SIMD); held at the one-block-per-entry kernel's 5 waves (96, no spills).
extern "C" __global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(5))) void computeForceBatchP(
    BatchArgs B, int ntotal) {
  __shared__ uint lq[FW_QBYTES / 4];
  batch_items(B, [&](int b) {
    const FParams_t* P;
    const Span S = batch_span(B, b, P);
    force_q(*P, S, ntotal, lq);
  });
}

// ------------------------------------------------------ neighbour statistics

// The largest number of bitmask words any in-range particle's stencil needs,
// on a state whose counts and offsets are current (after insert and scan; the
// particles need not be sorted): out[0] = max.  Sizes the reverse mode's tape.
__device__ __forceinline__ void max_words_impl(const FParams_t& P, const Span S, uint* out) {
  const int i = S.p0 + S.blk * (int)blockDim.x + (int)threadIdx.x;
  if (i >= S.p1) return;
  const int gc = pbuf<const int>(FPnts, FGCELL)[i];
  if (gc == FX_GRID_UNDEF) return;
  const int w = nbr_words(P, gc, pbuf<const uint>(FAccel, AGRIDOFF), pbuf<const uint>(FAccel, AGRIDCNT));
  atomicMax(out, (uint)w);
}

extern "C" __global__ void fxMaxWords(uint* out, int pnum) {
  max_words_impl(*cparams(&FParams), Span{0, pnum, 0, (int)blockIdx.x}, out);
}

extern "C" __global__ void fxMaxWordsBatch(BatchArgs B, uint* out) {
  const FParams_t* P;
  const Span S = batch_span(B, blockIdx.x, P);
  max_words_impl(*P, S, out);
}

// Sums over all in-range particles of the stencil candidates (members of the
// stencil cells, self included) and of the in-radius pairs (0 < d2 < rd2),
// on the cell-sorted set: out[0] += C, out[1] += P.  The measurement behind
// the flop model of the density and force stages; the role of the
// reference's neighbour-count validation (fluids4.0/fluid_system.cpp:959-1011).
__device__ __forceinline__ void count_pairs_impl(const FParams_t& P, const Span S, unsigned long long* out) {
  const int i = S.p0 + S.blk * (int)blockDim.x + (int)threadIdx.x;
  unsigned long long nc = 0, np = 0;
  if (i < S.p1) {
    const int gc = pbuf<const int>(FPnts, FGCELL)[i];
    if (gc != FX_GRID_UNDEF) {
      const fx_float3* __restrict__ pos = pbuf<const fx_float3>(FPnts, FPOS);
      const uint* __restrict__ off = pbuf<const uint>(FAccel, AGRIDOFF);
      const uint* __restrict__ cnt = pbuf<const uint>(FAccel, AGRIDCNT);
      const fx_float3 p = pos[i];
      const int base = stencil_base(P, gc);
      for (int r = 0; r < P.gridSrch * P.gridSrch; r++) {
        int jb, je;
        row_range(P, base, r, off, cnt, jb, je);
        nc += (unsigned long long)(je - jb);
        for (int j = jb; j < je; j++) {
          const fx_float3 q = pos[j];
          const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
          const float dsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
          np += (dsq < P.rd2 && dsq > 0.0f) ? 1u : 0u;
        }
      }
    }
  }
  // wave64 reduction, one atomic per wave
  for (int o = 32; o > 0; o >>= 1) {
    nc += __shfl_xor(nc, o, 64);
    np += __shfl_xor(np, o, 64);
  }
  if ((threadIdx.x & 63) == 0 && (nc | np) != 0) {
    atomicAdd(&out[0], nc);
    atomicAdd(&out[1], np);
  }
}

extern "C" __global__ void fxCountPairs(unsigned long long* out, int pnum) {
  count_pairs_impl(*cparams(&FParams), Span{0, pnum, 0, (int)blockIdx.x}, out);
}

extern "C" __global__ void fxCountPairsBatch(BatchArgs B, unsigned long long* out) {
  const FParams_t* P;
  const Span S = batch_span(B, blockIdx.x, P);
  count_pairs_impl(*P, S, out);
}

// ------------------------------------------------------ volume sampling

// Gaussian splat of the particles onto a brick of res.x * res.y * res.z
// voxels spanning [bmin, bmax] (sampleParticles,
// fluids4.0/fluid_kernels.cu:259-300), a consumer of the cell table: voxel
// (x, y, z) sits at bmin + (x/res.x, y/res.y, z/res.z) * (bmax - bmin) and
// sums gausskern * exp(-r^2 / h2), h2 = 2 r2 / 8, over the particles of its
// 27 stencil cells with 0 < d^2 < rd2 (r in simulation units), in the
// stencil's cell order; the result, times `scalar`, goes to
// brick[(z * res.y + y) * res.x + x].  Voxels whose cell is outside
// [1, gridRes - gridSrch] get 0 at that same index (the reference writes
// those at (y * res.z + z) * res.x + x, a transposed index).  Positions are
// the current set's; the cell table is the one of the last sort.
extern "C" __global__ void sampleParticles(float* brick, fx_int3 res, fx_float3 bmin, fx_float3 bmax, int numPnts,
                                           float scalar) {
  const FParams_t& P = *cparams(&FParams);
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z * blockDim.z + threadIdx.z;
  if (x >= res.x || y >= res.y || z >= res.z) return;
  const size_t out = ((size_t)z * res.y + y) * res.x + x;
  const float px = bmin.x + ((float)x / (float)res.x) * (bmax.x - bmin.x);
  const float py = bmin.y + ((float)y / (float)res.y) * (bmax.y - bmin.y);
  const float pz = bmin.z + ((float)z / (float)res.z) * (bmax.z - bmin.z);
  const int gx = (int)((px - P.gridMin.x) * P.gridDelta.x);
  const int gy = (int)((py - P.gridMin.y) * P.gridDelta.y);
  const int gz = (int)((pz - P.gridMin.z) * P.gridDelta.z);
  if (gx < 1 || gx > P.gridRes.x - P.gridSrch || gy < 1 || gy > P.gridRes.y - P.gridSrch || gz < 1 ||
      gz > P.gridRes.z - P.gridSrch) {
    brick[out] = 0.0f;
    return;
  }
  const int nadj = (P.gridRes.z + 1) * P.gridRes.x + 1;
  const int gs = (gy * P.gridRes.z + gz) * P.gridRes.x + gx - nadj;
  const fx_float3* __restrict__ pos = pbuf<const fx_float3>(FPnts, FPOS);
  const uint* __restrict__ off = pbuf<const uint>(FAccel, AGRIDOFF);
  const uint* __restrict__ cnt = pbuf<const uint>(FAccel, AGRIDCNT);
  const float h2 = 2.0f * P.r2 / 8.0f;
  float v = 0.0f;
  for (int c = 0; c < P.gridAdjCnt; c++) {
    const int cell = gs + P.gridAdj[c];
    const int j0 = (int)off[cell], j1 = j0 + (int)cnt[cell];
    for (int j = j0; j < j1 && j < numPnts; j++) {
      const fx_float3 q = pos[j];
      const float dx = px - q.x, dy = py - q.y, dz = pz - q.z;
      const float dsq = dx * dx + dy * dy + dz * dz;
      if (dsq < P.rd2 && dsq > 0.0f) {
        const float r = sqrtf(dsq * P.d2);
        v += P.gausskern * expf(-(r * r) / h2);
      }
    }
  }
  brick[out] = v * scalar;
}

// ------------------------------------------------------------------- emit

// emitParticles (fluids4.0/fluid_kernels.cu:231-257): each of `emit` threads
// respawns one particle at a point of the box (-170,10,-20)..(-190,60,20)
// quantized to 1/256 per axis by three bytes of a 32-bit value v, with
// velocity (5,-2,0) and zero veval and force.  The reference draws v from a
// per-particle XORWOW state and respawns particle v & (numPnts-1), so two
// threads can pick the same particle and the result depends on their order.
// Here v is a counter-based SplitMix64 of (frame, thread index) and the
// particles are a frame-seeded bijection of the thread index,
// j = (A i + b) mod numPnts with A prime (2654435761): min(emit, numPnts)
// distinct particles, whatever the thread order.
__device__ __forceinline__ unsigned long long fx_splitmix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

extern "C" __global__ void emitParticles(float frame, int emit, int numPnts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= emit || i >= numPnts) return;
  const unsigned long long h = fx_splitmix64((unsigned long long)__float_as_uint(frame));
  const uint v = (uint)(fx_splitmix64(h ^ fx_splitmix64((unsigned long long)i)) >> 32);
  const unsigned long long n = (unsigned long long)numPnts;
  const uint j = (uint)((2654435761ull * (unsigned long long)i + h % n) % n);
  const fx_float3 bmin = {-170.0f, 10.0f, -20.0f}, bmax = {-190.0f, 60.0f, 20.0f};
  const float ux = (float)(v & 0xFFu) / 256.0f;
  const float uy = (float)((v >> 8) & 0xFFu) / 256.0f;
  const float uz = (float)((v >> 16) & 0xFFu) / 256.0f;
  fx_float3 p;
  p.x = bmin.x + ux * (bmax.x - bmin.x);
  p.y = bmin.y + uy * (bmax.y - bmin.y);
  p.z = bmin.z + uz * (bmax.z - bmin.z);
  pbuf<fx_float3>(FPnts, FPOS)[j] = p;
  pbuf<fx_float3>(FPnts, FVEVAL)[j] = fx_float3{0.0f, 0.0f, 0.0f};
  pbuf<fx_float3>(FPnts, FVEL)[j] = fx_float3{5.0f, -2.0f, 0.0f};
  pbuf<fx_float3>(FPnts, FFORCE)[j] = fx_float3{0.0f, 0.0f, 0.0f};
}

// ----------------------------------------------------------------- advance

// Integrator (particles.cu:237-322): penalty walls on bound_min/bound_max
// (Y-min with slope and friction, Y-max with friction, X-min moving wall whose
// stiffness term is scaled by 'wall', X-max, Z-min, Z-max), gravity, the
// acceleration limit, the velocity limit applied to the old velocity, and the
// leapfrog update.  GRID_UNDEF particles are parked at (-1000)^3 with zero
// velocity.
__device__ __forceinline__ void advance_one(const FParams_t& P, int i, float time, float dt, float ss) {
  fx_float3* __restrict__ posb = pbuf<fx_float3>(FPnts, FPOS);
  fx_float3* __restrict__ velb = pbuf<fx_float3>(FPnts, FVEL);
  if (pbuf<const int>(FPnts, FGCELL)[i] == FX_GRID_UNDEF) {
    posb[i] = fx_float3{-1000.0f, -1000.0f, -1000.0f};
    velb[i] = fx_float3{0.0f, 0.0f, 0.0f};
    return;
  }
  fx_float3* __restrict__ vevb = pbuf<fx_float3>(FPnts, FVEVAL);
  const fx_float3 pos = posb[i];
  fx_float3 vel = velb[i];
  const fx_float3 ve = vevb[i];
  const fx_float3 f = pbuf<const fx_float3>(FPnts, FFORCE)[i];
  const float m = P.pmass;
  float ax = f.x * m, ay = f.y * m, az = f.z * m;
  const float rad = P.pradius, stiff = P.bound_stiff, damp = P.bound_damp;
  const float slope = P.bound_slope, fric = P.bound_friction;
  const fx_float3 bmin = P.bound_min, bmax = P.bound_max;
  float diff, adj, nx, ny, nz;

  // Y-min, sloped floor with friction
  diff = rad - (pos.y - (bmin.y + (pos.x - bmin.x) * slope)) * ss;
  if (diff > FX_EPSILON) {
    nx = -slope;
    ny = 1.0f - slope;
    nz = 0.0f;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    nx *= adj; ny *= adj; nz *= adj;
    ax += nx - ve.x * fric;
    ay += ny - ve.y * fric;
    az += nz - ve.z * fric;
  }
  // Y-max with friction
  diff = rad - (bmax.y - pos.y) * ss;
  if (diff > FX_EPSILON) {
    nx = 0.0f; ny = -1.0f; nz = 0.0f;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    nx *= adj; ny *= adj; nz *= adj;
    ax += nx - ve.x * fric;
    ay += ny - ve.y * fric;
    az += nz - ve.z * fric;
  }
  // X-min, moving wall; its stiffness term is scaled by 'wall'
  const float sarg = time * P.bound_wall_freq + pos.z / 400.0f;
  // in double, as the reference (its 0.5 are double literals) and the oracle:
  // an FP32 restatement would round (s + 1) force / 2 once where the double
  // path may round twice
  const float wall = (float)(((double)fx_sinf(sarg) * 0.5 + 0.5) * (double)P.bound_wall_force);
  diff = rad - (pos.x - (bmin.x + wall)) * ss;
  if (diff > FX_EPSILON) {
    nx = 1.0f; ny = 0.0f; nz = 0.0f;
    adj = wall * stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    ax += nx * adj; ay += ny * adj; az += nz * adj;
  }
  // X-max
  diff = rad - (bmax.x - pos.x) * ss;
  if (diff > FX_EPSILON) {
    nx = -1.0f; ny = 0.0f; nz = 0.0f;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    ax += nx * adj; ay += ny * adj; az += nz * adj;
  }
  // Z-min
  diff = rad - (pos.z - bmin.z) * ss;
  if (diff > FX_EPSILON) {
    nx = 0.0f; ny = 0.0f; nz = 1.0f;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    ax += nx * adj; ay += ny * adj; az += nz * adj;
  }
  // Z-max
  diff = rad - (bmax.z - pos.z) * ss;
  if (diff > FX_EPSILON) {
    nx = 0.0f; ny = 0.0f; nz = -1.0f;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    ax += nx * adj; ay += ny * adj; az += nz * adj;
  }
  // gravity
  ax += P.gravity.x;
  ay += P.gravity.y;
  az += P.gravity.z;
  // acceleration limit
  float speed = ax * ax + ay * ay + az * az;
  if (speed > P.AL2) {
    const float s = P.AL / sqrtf(speed);
    ax *= s; ay *= s; az *= s;
  }
  // velocity limit, applied to the old velocity
  speed = vel.x * vel.x + vel.y * vel.y + vel.z * vel.z;
  if (speed > P.VL2) {
    const float s = P.VL / sqrtf(speed);
    vel.x *= s; vel.y *= s; vel.z *= s;
  }
  // leapfrog
  const float vnx = ax * dt + vel.x, vny = ay * dt + vel.y, vnz = az * dt + vel.z;
  vevb[i] = fx_float3{(vel.x + vnx) * 0.5f, (vel.y + vny) * 0.5f, (vel.z + vnz) * 0.5f};
  velb[i] = fx_float3{vnx, vny, vnz};
  const float dts = dt / ss;
  posb[i] = fx_float3{pos.x + vnx * dts, pos.y + vny * dts, pos.z + vnz * dts};
}

extern "C" __global__ void advanceParticles(float time, float dt, float ss, int numPnts) {
  const int ppt = ppt_of(5);
  if (ppt == 1) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < numPnts) advance_one(*cparams(&FParams), i, time, dt, ss);
    return;
  }
  for (int k = 0; k < ppt; k++) {
    const int i = (blockIdx.x * ppt + k) * blockDim.x + threadIdx.x;
    if (i < numPnts) advance_one(*cparams(&FParams), i, time, dt, ss);
  }
}

// Same integrator with the simulation time read from device memory, for
// captured step graphs whose launches cannot carry a per-step time argument.
// fxPublishTables moves that clock on by dt each step (the host's
// AdvanceTime, particles.cpp:396-400).
extern "C" __global__ void advanceParticlesClock(const float* clock, float dt, float ss, int numPnts) {
  const int ppt = ppt_of(5);
  if (ppt == 1) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < numPnts) advance_one(*cparams(&FParams), i, clock[0], dt, ss);
    return;
  }
  for (int k = 0; k < ppt; k++) {
    const int i = (blockIdx.x * ppt + k) * blockDim.x + threadIdx.x;
    if (i < numPnts) advance_one(*cparams(&FParams), i, clock[0], dt, ss);
  }
}

// The set exchange of the counting sort (one launch, 64 threads): FPnts and
// FPntTmp receive the two 128-byte tables of bank[parity] and
// bank[parity ^ 1].  In a captured step it also moves the step clock: clock[1]
// holds the time of the step about to run, which becomes clock[0] (read by
// the advance of this step) while clock[1] moves on by dt.
extern "C" __global__ void fxPublishTables(const uint* __restrict__ bank, int parity, uint* __restrict__ pnts,
                                           uint* __restrict__ tmp, float* __restrict__ clock, float dt) {
  const int t = (int)threadIdx.x;
  constexpr int W = (int)(sizeof(cuDataX) / sizeof(uint));
  static_assert(2 * W <= 64, "one thread per word of the two tables");
  if (t < W) pnts[t] = bank[parity * W + t];
  else if (t < 2 * W) tmp[t - W] = bank[(parity ^ 1) * W + t - W];
  if (clock != nullptr && t == 0) {
    const float now = clock[1];
    clock[0] = now;
    clock[1] = now + dt;
  }
}

// Batch integrator: the time comes from clock[0] when clock is given, else
// from the argument; dt and the scale from each scene's parameters.
__device__ __forceinline__ void advance_entry(const BatchArgs& B, int b, const float* clock, float time) {
  const int ppt = ppt_of(5);
  const FParams_t* P;
  const Span S = batch_span(B, b, P);
  for (int k = 0; k < ppt; k++) {
    const int i = S.p0 + (S.blk * ppt + k) * (int)blockDim.x + (int)threadIdx.x;
    if (i < S.p1) advance_one(*P, i, clock != nullptr ? clock[0] : time, P->dt, P->sim_scale);
  }
}

extern "C" __global__ void advanceParticlesBatch(BatchArgs B, const float* clock, float time) {
  advance_entry(B, blockIdx.x, clock, time);
}

extern "C" __global__ void advanceParticlesBatchP(BatchArgs B, const float* clock, float time) {
  batch_items(B, [&](int b) { advance_entry(B, b, clock, time); });
}
