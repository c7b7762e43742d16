// adjoint.cpp -- reverse-mode rollouts of the SPH step (no reference
// counterpart): checkpointed forward, recomputation per segment, adjoint
// kernels of adjoint_gfx950.hip.
//
// A rollout of nsteps is cut into segments of `every` steps.  The forward
// saves the state (pos, vel, veval, clr, id, time) before every segment and
// after the last, and records the last segment's steps on the tape (the last
// two segments' when the tape has two buffers).  The backward walks the
// segments in reverse: a segment whose steps are not on the tape is
// recomputed from its checkpoint (the step is deterministic, so the
// recomputed steps are bit-identical to the forward's), recording each step
// right after its force stage; then the adjoint kernels run over the
// recorded steps in reverse.  The adjoint kernels run on a second stream, so
// that segment s-1 is recomputed (into a second tape buffer) while the
// adjoint of segment s runs.  The sorts of a rollout record their
// permutations (FSortSrc), through which the adjoint state moves from one
// step's slot order to the one before it.  For the length of a rollout the
// particles are also numbered by their slot at its start (see adjFrameIn).
// Memory: (segments
// + 1) x 44 B x N of checkpoints plus every x (60 B x N + 8 B x cells +
// W x 4 B x N of neighbour bitmask, W sized from the densest stencil) of tape, twice when there is
// more than one segment and the second buffer fits in 60 % of the free
// device memory.
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fx_common.hpp"
#include "particles.hpp"

namespace {

enum {
  ADJ_LOSS, ADJ_ADVANCE, ADJ_FORCE, ADJ_DENSITY, ADJ_REDUCE, ADJ_FRAME_IN, ADJ_FRAME_OUT, ADJ_UNFRAME, ADJ_TAPE,
  ADJ_CKPT, ADJ_NFUNC
};
const char* kAdjNames[ADJ_NFUNC] = {"adjLoss",    "adjAdvance",  "adjForce",   "adjDensity",    "adjReduce",
                                    "adjFrameIn", "adjFrameOut", "adjUnframe", "adjTapeRecord", "adjCheckpoint"};

// particle buffers of a checkpoint (CkptPtr in adjoint_gfx950.hip)
struct CkptPtr {
  void* v3[3];
  void* s1[2];
};

// source arrays of one tape slot (TapeSrc in adjoint_gfx950.hip)
struct TapeSrc {
  const void* v3[4];
  const void* s1[3];
  const void* cnt;
  const void* off;
};
const int kAdjThreads = 64;   // ADJ_THREADS: one wave per block
const int kRedThreads = 256;  // RED_THREADS

// kernel argument block (AdjArgs in adjoint_gfx950.hip)
struct AdjArgs {
  const void* params;
  const void* pbase;
  const void* btab;
  const void* pos;
  const void* vel;
  const void* veval;
  const void* force;
  const void* rho;
  const void* id;
  const void* gcell;
  const void* cnt;
  const void* off;
  const void* target;
  void* gx;
  void* gv;
  void* ge;
  void* gxp;
  void* gvp;
  void* gep;
  void* gF;
  void* gxs;
  void* ges;
  void* grho;
  void* partial;
  void* arec;
  void* drec;
  const void* bits;
  int nbits;
  float time;
  int accumulate;
  int ncap;
};

const size_t kCkptBytes = 44;   // pos, vel, veval (12 B each), clr, id (4 B each)
const size_t kTapeBytes = 60;   // pos, vel, veval, force (12 B each), rho, sort permutation, gcell (4 B each)
const size_t kVecFloats = 34;   // floats per particle of m_AdjVec
// plus, per recorded step, the density kernel's neighbour bitmask
// (m_AdjWords words per particle), written there directly by the forward

// one thread per particle, 256 per block
int launch_n(hipFunction_t f, const char* name, hipStream_t st, int n, void** args) {
  if (n <= 0) return FX_OK;
  FX_HIP(hipModuleLaunchKernel(f, (unsigned)((n + 255) / 256), 1, 1, 256, 1, 1, 0, st, args, nullptr), name);
  return FX_OK;
}

}  // namespace

void Particles::AdjRelease() {
  if (m_AdjSrc != nullptr && m_SortSrcSym != nullptr) {
    // a rollout that stopped on an error may have left the sort writing into
    // the buffer about to be freed
    void* null = nullptr;
    (void)hipStreamSynchronize(m_Stream);
    (void)hipMemcpyHtoD(m_SortSrcSym, &null, sizeof(null));
  }
  for (void* p : m_AdjAlloc) (void)hipFree(p);
  m_AdjAlloc.clear();
  m_AdjCkpt = m_AdjTape = nullptr;
  m_AdjBits = nullptr;
  m_AdjWordCaps = nullptr;
  m_AdjWords = FX_NBR_WORDS;
  m_AdjArec = m_AdjDrec = m_AdjTabs = nullptr;
  m_AdjVec = nullptr;
  m_AdjId0 = nullptr;
  m_AdjSrc = nullptr;
  m_AdjSrcPtr = nullptr;
  m_AdjPartial = m_AdjAccum = nullptr;
  m_AdjParams = m_AdjPBase = m_AdjBTab = m_AdjBFirst = nullptr;
  m_AdjBytes = 0;
  m_AdjLossDone = false;
  m_AdjBufSeg[0] = m_AdjBufSeg[1] = -1;
  m_AdjSteps = m_AdjEvery = m_AdjSegs = m_AdjN = m_AdjCells = 0;
}

int Particles::AdjConfigure(const char* path, int nsteps, int every) {
  if (mNumPoints <= 0) return fx::fail(FX_ESTATE, "adjoint: call Restart first");
  if (nsteps < 0) return fx::fail(FX_EINVAL, "adjoint: nsteps %d", nsteps);
  if (m_AdjModule == nullptr) {
    std::string p;
    if (path != nullptr) {
      p = path;
    } else {
      const size_t k = m_CodePath.find_last_of('/');
      p = (k == std::string::npos ? std::string() : m_CodePath.substr(0, k + 1)) + "adjoint_gfx950.hsaco";
    }
    FX_HIP(hipModuleLoad(&m_AdjModule, p.c_str()), "adjoint: loading the adjoint code object");
    for (int f = 0; f < ADJ_NFUNC; f++)
      FX_HIP(hipModuleGetFunction(&m_AdjFunc[f], m_AdjModule, kAdjNames[f]), kAdjNames[f]);
  }
  if (m_AdjStream == nullptr) {
    FX_HIP(hipStreamCreate(&m_AdjStream), "adjoint: stream");
    for (hipEvent_t& e : m_AdjEv) FX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming), "adjoint: events");
  }
  FX_HIP(hipStreamSynchronize(m_Stream), "adjoint");
  AdjRelease();
  const int n = mNumPoints;
  const size_t tape_step = kTapeBytes * (size_t)n + 8 * (size_t)m_ScanCells;
  // Bitmask words per particle a tape slot holds: twice the densest stencil
  // of the current state (counts and offsets from an insert and scan, which
  // reorder nothing), at least 32 and at most FX_NBR_WORDS, a multiple of 4;
  // FX_ADJOINT_TAPE_WORDS in the environment sets it.  A particle whose
  // stencil needs more during the rollout records no bits and is walked by
  // testing its candidates, forward and reverse.
  FX_TRY(InsertParticles());
  FX_TRY(PrefixScanParticles());
  int maxw = 0;
  FX_TRY(MaxNbrWords(&maxw));
  int words = 2 * maxw;
  words = words < 32 ? 32 : (words > FX_NBR_WORDS ? FX_NBR_WORDS : words);
  words = (words + 3) / 4 * 4;
  if (const char* w = getenv("FX_ADJOINT_TAPE_WORDS")) {
    const int v = atoi(w);
    if (v >= 1 && v <= FX_NBR_WORDS) words = v;
  }
  // memory budget: 60 % of the free device memory
  size_t free_b = 0, total_b = 0;
  FX_HIP(hipMemGetInfo(&free_b, &total_b), "adjoint: querying device memory");
  const double budget = 0.6 * (double)free_b;
  const double fixed = (double)(sizeof(float) * kVecFloats + 4 + 4 + 48 + 16) * n;
  const double per_step = (double)tape_step + 4.0 * words * n;
  if (every <= 0) {
    // The backward recomputes every segment but the last, (segments - 1) x
    // every steps: take the fewest segments whose tape (with its bitmasks)
    // and checkpoints fit the budget, of equal length.
    int longest = 1;  // longest segment that fits
    for (int e = nsteps > 1 ? nsteps : 1; e >= 1; e--) {
      const int sg = (nsteps + e - 1) / e;
      if (fixed + (double)kCkptBytes * n * (sg + 1) + per_step * (nsteps < e ? nsteps : e) <= budget) {
        longest = e;
        break;
      }
    }
    const int sg = nsteps > 0 ? (nsteps + longest - 1) / longest : 1;
    every = nsteps > 0 ? (nsteps + sg - 1) / sg : 1;
  }
  const int segs = nsteps == 0 ? 0 : (nsteps + every - 1) / every;
  const int seg_steps = nsteps < every ? nsteps : every;
  // two tape buffers when segments are recomputed and the second one fits
  // the budget: the backward then recomputes segment s-1 into one while the
  // adjoint of segment s reads the other
  // (FX_ADJOINT_TAPE_BUFFERS=1 in the environment keeps one buffer)
  const char* one = getenv("FX_ADJOINT_TAPE_BUFFERS");
  const bool two = segs > 1 && fixed + (double)kCkptBytes * n * (segs + 1) + 2.0 * per_step * seg_steps <= budget &&
                   !(one != nullptr && one[0] == '1');
  const int bufs = two ? 2 : 1;
  const int tape_steps = bufs * seg_steps;
  const int scenes = NumScenes();
  int blocks = 0;
  for (int s = 0; s < scenes; s++) blocks += (m_PBase[s + 1] - m_PBase[s] + kAdjThreads - 1) / kAdjThreads;
  auto alloc = [&](size_t bytes, void** out) -> int {
    if (bytes == 0) bytes = 8;
    FX_HIP(hipMalloc(out, bytes), "adjoint: allocating the tape and checkpoints");
    m_AdjAlloc.push_back(*out);
    m_AdjBytes += (int64_t)bytes;
    return FX_OK;
  };
  void* p;
  FX_TRY(alloc(kCkptBytes * (size_t)n * (size_t)(segs + 1), &p));
  m_AdjCkpt = static_cast<char*>(p);
  FX_TRY(alloc(tape_step * (size_t)tape_steps, &p));
  m_AdjTape = static_cast<char*>(p);
  FX_TRY(alloc(sizeof(float) * (size_t)n * kVecFloats, &p));
  m_AdjVec = static_cast<float*>(p);
  FX_TRY(alloc(sizeof(uint32_t) * (size_t)n, &p));
  m_AdjId0 = static_cast<uint32_t*>(p);
  FX_TRY(alloc(sizeof(uint32_t) * (size_t)n, &p));
  m_AdjSrc = static_cast<uint32_t*>(p);
  FX_TRY(alloc(2 * sizeof(void*), &p));
  m_AdjSrcPtr = static_cast<void**>(p);
  {
    void* ptrs[2] = {m_AdjSrc, nullptr};
    FX_HIP(hipMemcpyHtoD(m_AdjSrcPtr, ptrs, sizeof(ptrs)), "adjoint");
  }
  FX_TRY(alloc(sizeof(uint32_t) * (size_t)words * (size_t)n * (size_t)tape_steps, &p));
  m_AdjBits = static_cast<char*>(p);
  FX_TRY(alloc(2 * sizeof(int), &p));
  m_AdjWordCaps = static_cast<int*>(p);
  {
    const int caps[2] = {FX_NBR_WORDS, words};
    FX_HIP(hipMemcpyHtoD(m_AdjWordCaps, const_cast<int*>(caps), sizeof(caps)), "adjoint");
  }
  m_AdjWords = words;
  FX_TRY(alloc(48 * (size_t)n, &m_AdjArec));
  FX_TRY(alloc(16 * (size_t)n, &m_AdjDrec));
  FX_TRY(alloc(sizeof(cuDataX) * (size_t)(tape_steps + 1), &m_AdjTabs));
  {
    // FAccel tables: [0] the system's own, [1 + k] with ANBRBITS in tape slot k
    std::vector<cuDataX> tabs(tape_steps + 1);
    m_Accel.FillTable(&tabs[0]);
    for (int k = 0; k < tape_steps; k++) {
      tabs[k + 1] = tabs[0];
      tabs[k + 1].mbuf[ANBRBITS] =
          (uint64_t)(uintptr_t)(m_AdjBits + sizeof(uint32_t) * (size_t)words * (size_t)n * (size_t)k);
    }
    FX_HIP(hipMemcpyHtoD(m_AdjTabs, tabs.data(), sizeof(cuDataX) * tabs.size()), "adjoint");
  }
  FX_TRY(alloc(sizeof(double) * 3 * (size_t)(blocks > 0 ? blocks : 1), &p));
  m_AdjPartial = static_cast<double*>(p);
  FX_TRY(alloc(sizeof(double) * 3 * (size_t)scenes, &p));
  m_AdjAccum = static_cast<double*>(p);
  FX_TRY(alloc(sizeof(FParams_t) * (size_t)scenes, &m_AdjParams));
  FX_TRY(alloc(sizeof(int) * (size_t)(scenes + 1), &m_AdjPBase));
  FX_TRY(alloc(sizeof(int) * 2 * (size_t)(blocks > 0 ? blocks : 1), &m_AdjBTab));
  FX_TRY(alloc(sizeof(int) * (size_t)(scenes + 1), &m_AdjBFirst));
  std::vector<int> tab, first(scenes + 1, 0);
  for (int s = 0; s < scenes; s++) {
    first[s] = (int)tab.size() / 2;
    const int nb = (m_PBase[s + 1] - m_PBase[s] + kAdjThreads - 1) / kAdjThreads;
    for (int b = 0; b < nb; b++) {
      tab.push_back(s);
      tab.push_back(b);
    }
  }
  first[scenes] = (int)tab.size() / 2;
  if (!tab.empty()) FX_HIP(hipMemcpyHtoD(m_AdjBTab, tab.data(), tab.size() * sizeof(int)), "adjoint");
  FX_HIP(hipMemcpyHtoD(m_AdjBFirst, first.data(), first.size() * sizeof(int)), "adjoint");
  FX_HIP(hipMemcpyHtoD(m_AdjPBase, m_PBase.data(), sizeof(int) * (size_t)(scenes + 1)), "adjoint");
  // target defaults to zero
  FX_HIP(hipMemset(m_AdjVec + (size_t)n * 19, 0, sizeof(float) * 3 * (size_t)n), "adjoint");
  m_AdjSteps = nsteps;
  m_AdjEvery = every;
  m_AdjSegs = segs;
  m_AdjN = n;
  m_AdjCells = m_ScanCells;
  m_AdjBlocks = blocks;
  m_AdjScenes = scenes;
  m_AdjCkptTime.assign(segs + 1, 0.0f);
  m_AdjTapeTime.assign(tape_steps, 0.0f);
  m_AdjBufs = bufs;
  m_AdjTapeSteps = seg_steps;
  return FX_OK;
}

int Particles::AdjTables() {
  if (m_AdjN == 0 || m_AdjN != mNumPoints || m_AdjCells != m_ScanCells || m_AdjScenes != NumScenes())
    return fx::fail(FX_ESTATE, "adjoint: configure after Restart (the particle or cell count changed)");
  if (m_Batch) {
    FX_HIP(hipMemcpyHtoD(m_AdjParams, m_SceneParams.data(), sizeof(FParams_t) * m_SceneParams.size()),
           "adjoint");
  } else {
    FX_HIP(hipMemcpyHtoD(m_AdjParams, &m_Params, sizeof(FParams_t)), "adjoint");
  }
  return FX_OK;
}

int Particles::AdjSetTarget(const float* target, int64_t count) {
  if (m_AdjN == 0) return fx::fail(FX_ESTATE, "adjoint: call fx_adjoint_configure first");
  if (target == nullptr || count != (int64_t)m_AdjN)
    return fx::fail(FX_EINVAL, "adjoint: the target holds one position per particle (%d)", m_AdjN);
  FX_HIP(hipStreamSynchronize(m_Stream), "adjoint");
  FX_HIP(hipMemcpyHtoD(m_AdjVec + (size_t)m_AdjN * 19, const_cast<float*>(target), sizeof(float) * 3 * (size_t)count),
         "adjoint: target upload");
  return FX_OK;
}

int Particles::AdjLaunch(int fid, void* args, hipStream_t st) {
  if (m_AdjBlocks == 0) return FX_OK;
  void* a[] = {args};
  FX_HIP(hipModuleLaunchKernel(m_AdjFunc[fid], (unsigned)m_AdjBlocks, 1, 1, kAdjThreads, 1, 1, 0, st, a, nullptr),
         kAdjNames[fid]);
  return FX_OK;
}

// pos, vel, veval, clr, id of the current set to (save) or from checkpoint k, one kernel
int Particles::AdjCheckpointCopy(int k, int save) {
  char* c = m_AdjCkpt + kCkptBytes * (size_t)m_AdjN * (size_t)k;
  CkptPtr p = {{m_Points.gpu(FPOS), m_Points.gpu(FVEL), m_Points.gpu(FVEVAL)}, {m_Points.gpu(FCLR), m_Points.gpu(FID)}};
  int nn = m_AdjN;
  void* a[] = {&p, &c, &nn, &save};
  return launch_n(m_AdjFunc[ADJ_CKPT], kAdjNames[ADJ_CKPT], m_Stream, nn, a);
}

int Particles::AdjSaveCheckpoint(int k) {
  FX_TRY(AdjCheckpointCopy(k, 1));
  m_AdjCkptTime[k] = m_Params.time;
  return FX_OK;
}

int Particles::AdjRestoreCheckpoint(int k) {
  FX_TRY(AdjCheckpointCopy(k, 0));
  m_Params.time = m_AdjCkptTime[k];
  return FX_OK;
}

// One step (Particles::Run order); with record_slot >= 0 the step's sorted
// state after the force stage goes to that tape slot.
int Particles::AdjStep(int record_slot) {
  FX_TRY(InsertParticles());
  FX_TRY(PrefixScanParticles());
  FX_TRY(CountingSort());
  char* tabs = static_cast<char*>(m_AdjTabs);
  if (record_slot >= 0) {  // the density kernel writes this step's bitmask into the tape, m_AdjWords a particle
    FX_HIP(hipMemcpyAsync(m_NbrWordsSym, m_AdjWordCaps + 1, sizeof(int), hipMemcpyDeviceToDevice, m_Stream), "tape");
    FX_HIP(hipMemcpyAsync(m_Accel.SymbolAddress(), tabs + sizeof(cuDataX) * (size_t)(record_slot + 1),
                          sizeof(cuDataX), hipMemcpyDeviceToDevice, m_Stream), "tape");
  }
  FX_TRY(ComputePressure());
  FX_TRY(ComputeForce());
  if (record_slot >= 0) {
    FX_HIP(hipMemcpyAsync(m_Accel.SymbolAddress(), tabs, sizeof(cuDataX), hipMemcpyDeviceToDevice, m_Stream),
           "tape");
    FX_HIP(hipMemcpyAsync(m_NbrWordsSym, m_AdjWordCaps, sizeof(int), hipMemcpyDeviceToDevice, m_Stream), "tape");
    // pos, vel, veval, force, rho, id, gcell, then the cell counts and offsets: one kernel
    const size_t n = (size_t)m_AdjN;
    char* t = m_AdjTape + (kTapeBytes * n + 8 * (size_t)m_AdjCells) * (size_t)record_slot;
    TapeSrc src = {{m_Points.gpu(FPOS), m_Points.gpu(FVEL), m_Points.gpu(FVEVAL), m_Points.gpu(FFORCE)},
                   {m_Points.gpu(FPRESS), m_AdjSrc, m_Points.gpu(FGCELL)},
                   m_Accel.gpu(AGRIDCNT),
                   m_Accel.gpu(AGRIDOFF)};
    int nn = m_AdjN, cells = m_AdjCells;
    void* a[] = {&src, &t, &nn, &cells};
    FX_TRY(launch_n(m_AdjFunc[ADJ_TAPE], kAdjNames[ADJ_TAPE], m_Stream, nn > cells ? nn : cells, a));
    m_AdjTapeTime[record_slot] = m_Params.time;
  }
  FX_TRY(Advance());
  AdvanceTime();
  return FX_OK;
}

int Particles::AdjSortSrc(bool on) {
  // the sorts of a rollout write their permutation into m_AdjSrc
  FX_HIP(hipMemcpyAsync(m_SortSrcSym, m_AdjSrcPtr + (on ? 0 : 1), sizeof(void*), hipMemcpyDeviceToDevice, m_Stream),
         "adjoint");
  return FX_OK;
}

int Particles::AdjForward(double* loss, double* ms) {
  FX_TRY(AdjTables());
  ReleaseGraph();
  fx::Event e0, e1;
  if (ms) {
    FX_HIP(e0.create(), "adjoint");
    FX_HIP(e1.create(), "adjoint");
    FX_HIP(hipEventRecord(e0.e, m_Stream), "adjoint");
  }
  {  // the rollout frame: FID := slot, the target follows
    void* fid = m_Points.gpu(FID);
    float* tgt = m_AdjVec + 19 * (size_t)m_AdjN;
    float* tgf = m_AdjVec + 22 * (size_t)m_AdjN;
    int nn = m_AdjN;
    void* a[] = {&fid, &m_AdjId0, &tgt, &tgf, &nn};
    FX_TRY(launch_n(m_AdjFunc[ADJ_FRAME_IN], kAdjNames[ADJ_FRAME_IN], m_Stream, nn, a));
  }
  // The last segment is recorded, and with two tape buffers the one before
  // it too: the backward then finds both in the tape and recomputes only the
  // segments before them.
  const int last = m_AdjSegs - 1;
  const int first_rec = m_AdjBufs == 2 ? last - 1 : last;
  m_AdjBufSeg[0] = m_AdjBufSeg[1] = -1;
  FX_TRY(AdjSortSrc(true));
  for (int t = 0; t < m_AdjSteps; t++) {
    const int seg = t / m_AdjEvery;
    if (t % m_AdjEvery == 0) FX_TRY(AdjSaveCheckpoint(seg));
    FX_TRY(AdjStep(seg >= first_rec ? (seg % m_AdjBufs) * m_AdjTapeSteps + t - seg * m_AdjEvery : -1));
  }
  for (int sg = first_rec < 0 ? 0 : first_rec; sg <= last; sg++) m_AdjBufSeg[sg % m_AdjBufs] = sg;
  FX_TRY(AdjSortSrc(false));
  FX_TRY(AdjSaveCheckpoint(m_AdjSegs));
  // loss and seed on the final state
  const size_t n = (size_t)m_AdjN;
  float* V = m_AdjVec;
  AdjArgs A = {};
  A.params = m_AdjParams;
  A.pbase = m_AdjPBase;
  A.btab = m_AdjBTab;
  A.pos = m_Points.gpu(FPOS);
  A.id = m_Points.gpu(FID);
  A.target = V + 22 * n;
  A.gx = V;
  A.gv = V + 3 * n;
  A.ge = V + 6 * n;
  A.partial = m_AdjPartial;
  FX_TRY(AdjLaunch(ADJ_LOSS, &A, m_Stream));
  int k0 = 0, k1 = 1, add = 0;
  void* rargs[] = {&m_AdjPartial, &m_AdjBFirst, &m_AdjAccum, &k0, &k1, &add};
  FX_HIP(hipModuleLaunchKernel(m_AdjFunc[ADJ_REDUCE], (unsigned)m_AdjScenes, 1, 1, kRedThreads, 1, 1, 0, m_Stream,
                               rargs, nullptr),
         "adjReduce");
  {  // the system leaves the call with its original ids
    void* fid = m_Points.gpu(FID);
    int nn = m_AdjN;
    void* a[] = {&fid, &m_AdjId0, &nn};
    FX_TRY(launch_n(m_AdjFunc[ADJ_FRAME_OUT], kAdjNames[ADJ_FRAME_OUT], m_Stream, nn, a));
  }
  if (ms) FX_HIP(hipEventRecord(e1.e, m_Stream), "adjoint");
  std::vector<double> acc(3 * (size_t)m_AdjScenes);
  FX_HIP(hipMemcpyAsync(acc.data(), m_AdjAccum, acc.size() * sizeof(double), hipMemcpyDeviceToHost, m_Stream),
         "adjoint");
  FX_HIP(hipStreamSynchronize(m_Stream), "adjoint");
  for (int s = 0; s < m_AdjScenes; s++) {
    const int ns = m_PBase[s + 1] - m_PBase[s];
    if (loss) loss[s] = ns > 0 ? acc[3 * s] / (double)ns : 0.0;
  }
  if (ms) {
    float f = 0.0f;
    FX_HIP(hipEventElapsedTime(&f, e0.e, e1.e), "adjoint");
    *ms = f;
  }
  m_AdjLossDone = true;
  return FX_OK;
}

// Segment seg recomputed from its checkpoint on the system stream and
// recorded into tape buffer seg % m_AdjBufs, once the adjoint that last read
// that buffer has finished.
int Particles::AdjRecompute(int seg) {
  const int b = seg % m_AdjBufs;
  FX_HIP(hipStreamWaitEvent(m_Stream, m_AdjEv[2 + b], 0), "adjoint");
  FX_TRY(AdjRestoreCheckpoint(seg));
  const int s0 = seg * m_AdjEvery;
  const int s1 = (s0 + m_AdjEvery < m_AdjSteps) ? s0 + m_AdjEvery : m_AdjSteps;
  m_AdjBufSeg[b] = -1;
  for (int t = s0; t < s1; t++) FX_TRY(AdjStep(b * m_AdjTapeSteps + t - s0));
  FX_HIP(hipEventRecord(m_AdjEv[b], m_Stream), "adjoint");
  m_AdjBufSeg[b] = seg;
  return FX_OK;
}

int Particles::AdjBackward(float* gpos, float* gvel, float* gveval, double* gvisc, double* gk, double* ms) {
  if (!m_AdjLossDone) return fx::fail(FX_ESTATE, "adjoint: call fx_adjoint_forward before fx_adjoint_backward");
  FX_TRY(AdjTables());
  ReleaseGraph();
  fx::Event e0, e1;
  if (ms) {
    FX_HIP(e0.create(), "adjoint");
    FX_HIP(e1.create(), "adjoint");
    FX_HIP(hipEventRecord(e0.e, m_Stream), "adjoint");
  }
  const size_t n = (size_t)m_AdjN;
  float* V = m_AdjVec;
  const size_t tape_step = kTapeBytes * n + 8 * (size_t)m_AdjCells;
  // The adjoint kernels run on their own stream; the system stream
  // recomputes segment s-1 into the other tape buffer meanwhile.  Events:
  // [b] buffer b holds its segment, [2 + b] the adjoint reading buffer b is
  // done, [4] join.
  FX_HIP(hipEventRecord(m_AdjEv[4], m_Stream), "adjoint");
  FX_HIP(hipStreamWaitEvent(m_AdjStream, m_AdjEv[4], 0), "adjoint");
  for (int b = 0; b < 2; b++) {
    FX_HIP(hipEventRecord(m_AdjEv[2 + b], m_AdjStream), "adjoint");
    FX_HIP(hipEventRecord(m_AdjEv[b], m_Stream), "adjoint");
  }
  bool first = true;
  FX_TRY(AdjSortSrc(true));
  // the adjoint state alternates between two sets: [cur] holds that of the
  // step's outputs, the step writes that of its inputs into [cur ^ 1]
  float* set[2] = {V, V + 25 * n};
  int cur = 0;
  auto held = [&](int sg) { return m_AdjBufSeg[sg % m_AdjBufs] == sg; };
  if (m_AdjSegs > 0 && !held(m_AdjSegs - 1)) FX_TRY(AdjRecompute(m_AdjSegs - 1));
  for (int seg = m_AdjSegs - 1; seg >= 0; seg--) {
    // beside the adjoint of this segment, unless the forward recorded it
    if (seg > 0 && m_AdjBufs == 2 && !held(seg - 1)) FX_TRY(AdjRecompute(seg - 1));
    const int b = seg % m_AdjBufs;
    FX_HIP(hipStreamWaitEvent(m_AdjStream, m_AdjEv[b], 0), "adjoint");
    const int s0 = seg * m_AdjEvery;
    const int s1 = (s0 + m_AdjEvery < m_AdjSteps) ? s0 + m_AdjEvery : m_AdjSteps;
    for (int t = s1 - 1; t >= s0; t--) {
      const int k = b * m_AdjTapeSteps + t - s0;
      const char* T = m_AdjTape + tape_step * (size_t)k;
      AdjArgs A = {};
      A.params = m_AdjParams;
      A.pbase = m_AdjPBase;
      A.btab = m_AdjBTab;
      A.pos = T;
      A.vel = T + 12 * n;
      A.veval = T + 24 * n;
      A.force = T + 36 * n;
      A.rho = T + 48 * n;
      A.id = T + 52 * n;
      A.gcell = T + 56 * n;
      A.cnt = T + 60 * n;
      A.off = T + 60 * n + 4 * (size_t)m_AdjCells;
      A.target = V + 22 * n;
      A.gx = set[cur];
      A.gv = set[cur] + 3 * n;
      A.ge = set[cur] + 6 * n;
      A.gxp = set[cur ^ 1];
      A.gvp = set[cur ^ 1] + 3 * n;
      A.gep = set[cur ^ 1] + 6 * n;
      A.gF = V + 9 * n;
      A.gxs = V + 12 * n;
      A.ges = V + 15 * n;
      A.grho = V + 18 * n;
      A.partial = m_AdjPartial;
      A.arec = m_AdjArec;
      A.drec = m_AdjDrec;
      A.bits = m_AdjBits + sizeof(uint32_t) * (size_t)m_AdjWords * n * (size_t)k;
      A.nbits = (int)n;
      A.ncap = m_AdjWords;
      A.time = m_AdjTapeTime[k];
      A.accumulate = first ? 0 : 1;
      FX_TRY(AdjLaunch(ADJ_ADVANCE, &A, m_AdjStream));
      FX_TRY(AdjLaunch(ADJ_FORCE, &A, m_AdjStream));
      FX_TRY(AdjLaunch(ADJ_DENSITY, &A, m_AdjStream));
      first = false;
      cur ^= 1;
    }
    FX_HIP(hipEventRecord(m_AdjEv[2 + b], m_AdjStream), "adjoint");
    if (seg > 0 && m_AdjBufs == 1) FX_TRY(AdjRecompute(seg - 1));  // one buffer: after its adjoint
  }
  FX_HIP(hipEventRecord(m_AdjEv[4], m_AdjStream), "adjoint");
  FX_HIP(hipStreamWaitEvent(m_Stream, m_AdjEv[4], 0), "adjoint");
  FX_TRY(AdjSortSrc(false));
  // the system continues from the end of the rollout, with its original ids;
  // the gradients go back to original-id order (into the gF..ges scratch)
  FX_TRY(AdjRestoreCheckpoint(m_AdjSegs));
  {
    void* fid = m_Points.gpu(FID);
    float* out = V + 9 * n;
    int nn = m_AdjN;
    void* a[] = {&fid, &m_AdjId0, &nn};
    FX_TRY(launch_n(m_AdjFunc[ADJ_FRAME_OUT], kAdjNames[ADJ_FRAME_OUT], m_Stream, nn, a));
    // the adjoint of the rollout's inputs, in its first step's order before the sort
    float* g = set[cur];
    void* b[] = {&g, &out, &m_AdjId0, &nn};
    FX_TRY(launch_n(m_AdjFunc[ADJ_UNFRAME], kAdjNames[ADJ_UNFRAME], m_Stream, nn, b));
  }
  m_AdjLossDone = false;
  std::vector<double> acc(3 * (size_t)m_AdjScenes, 0.0);
  if (!first) {
    int k0 = 1, k1 = 3, add = 0;
    void* rargs[] = {&m_AdjPartial, &m_AdjBFirst, &m_AdjAccum, &k0, &k1, &add};
    FX_HIP(hipModuleLaunchKernel(m_AdjFunc[ADJ_REDUCE], (unsigned)m_AdjScenes, 1, 1, kRedThreads, 1, 1, 0,
                                 m_Stream, rargs, nullptr),
           "adjReduce");
  }
  if (ms) FX_HIP(hipEventRecord(e1.e, m_Stream), "adjoint");
  if (!first)
    FX_HIP(hipMemcpyAsync(acc.data(), m_AdjAccum, acc.size() * sizeof(double), hipMemcpyDeviceToHost, m_Stream),
           "adjoint");
  const size_t vb = sizeof(float) * 3 * n;
  if (gpos) FX_HIP(hipMemcpyAsync(gpos, V + 9 * n, vb, hipMemcpyDeviceToHost, m_Stream), "adjoint");
  if (gvel) FX_HIP(hipMemcpyAsync(gvel, V + 12 * n, vb, hipMemcpyDeviceToHost, m_Stream), "adjoint");
  if (gveval) FX_HIP(hipMemcpyAsync(gveval, V + 15 * n, vb, hipMemcpyDeviceToHost, m_Stream), "adjoint");
  FX_HIP(hipStreamSynchronize(m_Stream), "adjoint");
  for (int s = 0; s < m_AdjScenes; s++) {
    if (gk) gk[s] = acc[3 * s + 1];
    if (gvisc) gvisc[s] = acc[3 * s + 2];
  }
  if (ms) {
    float f = 0.0f;
    FX_HIP(hipEventElapsedTime(&f, e0.e, e1.e), "adjoint");
    *ms = f;
  }
  return FX_OK;
}
