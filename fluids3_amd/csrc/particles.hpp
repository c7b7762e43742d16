// particles.hpp -- host orchestration of the SPH step on one MI355X.
//
// Mirrors the reference's Particles class (fluids5.0/src/particles.h:29-96):
// same stage methods in the same order, same DataX sets (m_Points,
// m_PointsTemp, m_Accel), same FParams_t block, kernels and globals resolved
// by name from one code object.  Differences, all on the device-memory side:
//   - CopyAllToTemp's eight device copies (particles.cpp:698-701) become an
//     exchange of the two sets' device pointers plus a 2 x 128-byte table
//     update, so the counting sort still reads FPntTmp and writes FPnts;
//   - all work goes to one owned HIP stream, which can be captured into a
//     graph;
//   - errors are status codes.
#pragma once

#include <hip/hip_runtime.h>

#include "datax.hpp"
#include "fluidsx.h"

// kernel slots (particles.h:104-111 plus the additions of this module)
#define FUNC_INSERT 0
#define FUNC_COUNTING_SORT 1
#define FUNC_COMPUTE_PRESS 2
#define FUNC_COMPUTE_FORCE 3
#define FUNC_ADVANCE 4
#define FUNC_FPREFIXSUM 5
#define FUNC_FPREFIXFIXUP 6
#define FUNC_SORT_INDEX 7
#define FUNC_FILL 8
#define FUNC_ADVANCE_CLOCK 9
#define FUNC_PUBLISH 10  // fxPublishTables: the set exchange (and the step clock of a captured step)
#define FUNC_COUNT_PAIRS 11
#define FUNC_INSERT_B 12
#define FUNC_SORT_INDEX_B 13
#define FUNC_SORT_FULL_B 14
#define FUNC_PRESS_B 15
#define FUNC_FORCE_B 16
#define FUNC_ADVANCE_B 17
#define FUNC_COUNT_PAIRS_B 18
#define FUNC_SAMPLE 19
#define FUNC_EMIT 20
#define FUNC_MAXWORDS 21
#define FUNC_MAXWORDS_B 22
#define FUNC_MARK 23
#define FUNC_INSERT_BP 24  // persistent batch kernels
#define FUNC_SORT_INDEX_BP 25
#define FUNC_SORT_FULL_BP 26
#define FUNC_PRESS_BP 27
#define FUNC_FORCE_BP 28
#define FUNC_ADVANCE_BP 29
#define FUNC_SCAN_SUMS 30  // the step's two-launch scans
#define FX_SCAN2_THREADS 512  // threads per block of fxScanSums / fxScanApply
#define FX_SCAN2_BLOCK 2048   // elements per block (4 per thread)
#define FX_SCAN2_MAX_BLOCKS 8192
#define FUNC_SCAN_APPLY 31
#define FUNC_MAX 32

#include <map>
#include <string>
#include <vector>

// Kernel argument block of the *Batch kernels (BatchArgs in the kernel file).
struct FxBatchArgs {
  const void* params;
  const void* pbase;
  const void* cbase;
  const void* btab;
  void* queue;  // claim counters of a persistent launch, else null
  int nitems;   // block-table entries
  int pad;
};

class Particles {
 public:
  Particles();
  ~Particles();
  Particles(const Particles&) = delete;
  Particles& operator=(const Particles&) = delete;

  // set-up (particles.cpp:68-119, 267-300, 548-679)
  int Initialize(const char* code_object_path);
  int Restart(const fx_scene& scene);
  // Many independent scenes in one system: particles and cell tables are
  // concatenated, every stage is one launch over all scenes.
  int RestartBatch(const fx_scene* scenes, int count);
  int NumScenes() const { return m_Batch ? (int)m_SceneParams.size() : 1; }
  int SceneRange(int s, int* p0, int* np, int* c0, int* nc) const;
  int SceneParams(int s, FParams_t* out) const;
  int ReallocateParticles(int maxpnt);
  int RebuildAccelGrid();
  int UpdateParams();
  int UpdateGPUAccess();
  int SetParams(const FParams_t& p);
  int SetBlockSize(int threads);
  int SetStageBlockSize(int stage, int threads);
  int StageBlockSize(int stage) const { return m_StageThreads[stage]; }
  // particles per thread of the insert, sort and advance stages (1..8)
  int SetStagePpt(int stage, int ppt);
  int StagePpt(int stage) const { return m_StagePpt[stage]; }
  // persistent (fixed-grid, work-queue) launches of the batch stages
  int SetPersistent(bool on);
  bool Persistent() const { return m_Persistent; }

  // the step (particles.cpp:382-400)
  int Run();
  int InsertParticles();
  int PrefixScanParticles();
  int CountingSort();
  int ComputePressure();
  int ComputeForce();
  int Advance();
  void AdvanceTime();
  int Stage(int stage);

  // graphs
  int CaptureGraph();
  int LaunchGraph(int nlaunch);

  // transfers (particles.cpp:703-711)
  int CommitAll();
  int Upload(int buf, const void* host, int64_t count);  // FID: a whole permutation of 0..n-1 only
  int Download(int buf, void* host, int64_t count);
  static int CheckIds(const uint32_t* id, int64_t n, const char* who);
  int DownloadAccel(int buf, void* host, int64_t count);
  int Synchronize();
  int Scan(const uint32_t* in, uint32_t* out, int n);
  int ScanDeviceNoAlloc(const uint32_t* in, uint32_t* out, int n);
  int Mark(int tag);

  // timing
  int ProfileStages(int nsteps, double* ms_out);
  int TimeSteps(int nsteps, bool use_graph, double* ms_total);
  int BenchSteps(int nsteps, bool use_graph, double* ms_total, double* stage_ms);
  int CountPairs(int64_t* cand, int64_t* pairs);
  int MaxNbrWords(int* words);  // after insert and scan: the largest stencil's bitmask words
  int Autotune(int nsteps, int* chosen, double* stage_ms);  // autotune.cpp
  int SavePoints(const char* path);                         // points_io.cpp
  int SampleParticles(float* brick, const int* res, const float* bmin, const float* bmax, float scalar);
  int EmitParticles(float frame, int emit);
  int LoadPoints(const char* path);

  // reverse mode (adjoint.cpp): checkpointed rollouts and their gradients
  int AdjConfigure(const char* code_object_path, int nsteps, int every);
  int AdjSetTarget(const float* target, int64_t count);
  int AdjForward(double* loss, double* ms);
  int AdjBackward(float* gpos, float* gvel, float* gveval, double* gvisc, double* gk, double* ms);
  int64_t AdjMemory() const { return m_AdjBytes; }
  int AdjSegment() const { return m_AdjEvery; }

  int NumPoints() const { return mNumPoints; }
  int CellTableSize() const { return m_ScanCells; }
  const FParams_t& Params() const { return m_Params; }
  hipStream_t Stream() const { return m_Stream; }

 private:
  int Launch(int fid, unsigned blocks, unsigned threads, void** args);
  // one thread per particle, with the block size chosen for the stage
  int LaunchP(int fid, int stage, void** args);
  int ScanLevels(void* array1, void* scan1, void* array2, void* scan2, void* array3, void* scan3, int n);
  // exclusive scan in two launches (fxScanSums, fxScanApply); sums holds one
  // uint per block of 2048 elements
  // exclusive scans of up to two arrays (n2 = 0: one) in two launches
  int ScanTwoLaunch(void* in, void* out, void* sums, int n, void* in2 = nullptr, void* out2 = nullptr,
                    void* sums2 = nullptr, int n2 = 0);
  int PublishTables();
  int SetDeviceClock();
  void ReleaseGraph();
  int BuildBlockTables();
  // one batch launch of a stage: kernel fid, one block per block-table
  // entry, or its persistent variant fid_p (0: none) after resetting the
  // stage's claim counters
  int BatchLaunch(int fid, int fid_p, int stage, void* extra0, void* extra1, void* extra2);
  int ResidentBlocks(int fid, int threads, int* blocks);
  int UploadPpt();
  int FinishRestart(int n);
  // adjoint helpers
  void AdjRelease();
  int AdjTables();
  int AdjStep(int record_slot);
  int AdjCheckpointCopy(int k, int save);
  int AdjSaveCheckpoint(int k);
  int AdjRestoreCheckpoint(int k);
  int AdjLaunch(int fid, void* args, hipStream_t st);
  int AdjRecompute(int seg);
  int AdjSortSrc(bool on);

  hipModule_t m_Module = nullptr;
  hipFunction_t m_Func[FUNC_MAX] = {};
  hipStream_t m_Stream = nullptr;
  void* m_cuParams = nullptr;  // device address of FParams

  int mNumPoints = 0;
  int mMaxPoints = 0;
  int m_Frame = 0;
  DataX m_Points;      // current, cell-sorted set
  DataX m_PointsTemp;  // the other device set
  DataX m_Accel;
  FParams_t m_Params;

  // device copy of both sets' pointer tables: [0] = set A, [1] = set B
  void* m_Tables = nullptr;
  int m_Parity = 0;  // which set m_Points holds
  // the last stage (FX_STAGE_*) run on the device since the last insert, -1
  // after a restart; Stage() refuses a stage whose input tables it has not made
  int m_Done = -1;
  float* m_Clock = nullptr;  // device {time of the step being run, time of the next step}
  void* m_Stats = nullptr;  // two uint64 counters of fxCountPairs
  void* m_ScanScratch = nullptr;  // this device's scratch array (noalloc.hip), not owned
  void* m_NbrWordsSym = nullptr;  // the module's FNbrWords (bitmask words a particle may record)
  void* m_SortSrcSym = nullptr;   // the module's FSortSrc (where the sort writes its permutation, or null)

  hipGraph_t m_Graph = nullptr;
  hipGraphExec_t m_GraphExec = nullptr;
  bool m_Capturing = false;
  int m_GraphParity = 0;
  // threads per block of each stage's per-particle kernels (insert, prefix,
  // sort, pressure, force, advance): the sizes fx_autotune keeps choosing on
  // the dam-break and beach scenes at 4M particles
  int m_StageThreads[FX_NUM_STAGES] = {128, 256, 512, 256, 256, 64};
  int m_StagePpt[FX_NUM_STAGES] = {1, 1, 1, 1, 1, 1};
  void* m_PptSym = nullptr;     // module constant FPpt
  bool m_Persistent = false;    // batch stages run as persistent launches
  uint32_t* m_WorkQ = nullptr;  // 8 claim counters per stage
  int m_NumCU = 0;
  std::map<std::pair<int, int>, int> m_Resident;  // (function, threads) -> resident blocks on the device

  // batch of scenes (m_Batch); a single scene uses the reference-named kernels
  bool m_Batch = false;
  int m_ScanCells = 0;  // entries of the cell tables: gridTotal+1, or the sum over scenes
  std::vector<FParams_t> m_SceneParams;
  std::vector<int> m_PBase, m_CBase;  // nscene+1 entries each
  void* m_DParams = nullptr;          // FParams_t per scene
  void* m_DPBase = nullptr;
  void* m_DCBase = nullptr;
  std::map<int, std::pair<void*, int>> m_BTab;  // particles per block -> (int2 table, entries)

  // adjoint (reverse mode)
  std::string m_CodePath;
  hipModule_t m_AdjModule = nullptr;
  hipStream_t m_AdjStream = nullptr;     // the backward's adjoint kernels (default priority)
  hipEvent_t m_AdjEv[5] = {};            // per tape buffer: recomputed [0,1], adjoint done [2,3]; join [4]
  int m_AdjBufs = 1, m_AdjTapeSteps = 0;  // tape buffers (2 when segments are recomputed) and slots in each
  hipFunction_t m_AdjFunc[10] = {};
  int m_AdjSteps = 0, m_AdjEvery = 0, m_AdjSegs = 0, m_AdjN = 0, m_AdjCells = 0;
  bool m_AdjLossDone = false;
  int m_AdjBufSeg[2] = {-1, -1};  // the segment each tape buffer holds, -1 for none
  int64_t m_AdjBytes = 0;
  std::vector<void*> m_AdjAlloc;         // every device allocation of the adjoint
  std::vector<float> m_AdjCkptTime;      // checkpoint k = state before step k*every
  std::vector<float> m_AdjTapeTime;      // time of each recorded step of the segment
  char* m_AdjCkpt = nullptr;             // (segs+1) x 44 B x N
  char* m_AdjTape = nullptr;             // every x (60 B x N + 8 B x cells)
  char* m_AdjBits = nullptr;             // every x m_AdjWords x 4 B x N: each recorded step's bitmask
  int m_AdjWords = FX_NBR_WORDS;         // bitmask words a tape slot holds per particle
  int* m_AdjWordCaps = nullptr;          // device {FX_NBR_WORDS, m_AdjWords}: values copied into FNbrWords
  void* m_AdjArec = nullptr;             // 48 B x N pair records of the force adjoint
  void* m_AdjDrec = nullptr;             // 16 B x N pair records of the density adjoint
  void* m_AdjTabs = nullptr;             // FAccel tables routing ANBRBITS into the tape
  float* m_AdjVec = nullptr;             // gx, gv, ge, gF, gxs, ges (3N floats each), grho (N), target,
                                         // target in the rollout frame (3N each), a second gx, gv, ge
  uint32_t* m_AdjId0 = nullptr;          // original id of each slot at the rollout's start
  uint32_t* m_AdjSrc = nullptr;          // the last sort's permutation (the module's FSortSrc points here
                                         // during rollouts)
  void** m_AdjSrcPtr = nullptr;          // device {m_AdjSrc, nullptr}: values copied into FSortSrc
  double* m_AdjPartial = nullptr;        // 3 per block
  double* m_AdjAccum = nullptr;          // 3 per scene
  void* m_AdjParams = nullptr;           // FParams_t per scene
  void* m_AdjPBase = nullptr;
  void* m_AdjBTab = nullptr;             // int2 per block
  void* m_AdjBFirst = nullptr;           // first block of each scene (scenes + 1)
  int m_AdjBlocks = 0, m_AdjScenes = 0;
};
