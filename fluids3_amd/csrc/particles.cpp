// particles.cpp -- host side of the SPH step (behaviour of
// fluids5.0/src/particles.cpp: Initialize :68-93, Restart :95-119,
// ReallocateParticles :267-300, Run :382-394, RebuildAccelGrid :548-626,
// UpdateParams :657-679, stage methods :716-919).
#include "particles.hpp"

#include <cstring>
#include <vector>

#include "fx_common.hpp"
#include "fx_scene.hpp"
#include "noalloc.hpp"

#define SCAN_BLOCKSIZE 512
#define SCAN_ELEMS (2 * SCAN_BLOCKSIZE)

static const char* kKernelNames[FUNC_MAX] = {
    "insertParticles",   "countingSortFullRanked", "computePressure", "computeForce",
    "advanceParticles",  "prefixSum",        "prefixFixup",     "countingSortIndex",
    "fxFillUint",        "advanceParticlesClock", "fxPublishTables", "fxCountPairs",
    "insertParticlesBatch", "countingSortIndexBatch", "countingSortFullBatch", "computePressureBatch",
    "computeForceBatch", "advanceParticlesBatch", "fxCountPairsBatch", "sampleParticles",
    "emitParticles",     "fxMaxWords",       "fxMaxWordsBatch",  "fxMark",
    "insertParticlesBatchP", "countingSortIndexBatchP", "countingSortFullBatchP", "computePressureBatchP",
    "computeForceBatchP", "advanceParticlesBatchP", "fxScanSums", "fxScanApply"};

Particles::Particles() { memset(&m_Params, 0, sizeof(m_Params)); }

static void free_batch(void*& a, void*& b, void*& c, std::map<int, std::pair<void*, int>>& t) {
  for (void** p : {&a, &b, &c})
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  for (auto& e : t)
    if (e.second.first) (void)hipFree(e.second.first);
  t.clear();
}

Particles::~Particles() {
  ReleaseGraph();
  AdjRelease();
  for (hipEvent_t& e : m_AdjEv)
    if (e) (void)hipEventDestroy(e);
  if (m_AdjStream) (void)hipStreamDestroy(m_AdjStream);
  if (m_AdjModule) (void)hipModuleUnload(m_AdjModule);
  free_batch(m_DParams, m_DPBase, m_DCBase, m_BTab);
  m_Points.DeleteAllBuffers();
  m_PointsTemp.DeleteAllBuffers();
  m_Accel.DeleteAllBuffers();
  if (m_Tables) (void)hipFree(m_Tables);
  if (m_Clock) (void)hipFree(m_Clock);
  if (m_Stats) (void)hipFree(m_Stats);
  if (m_WorkQ) (void)hipFree(m_WorkQ);
  if (m_Stream) (void)hipStreamDestroy(m_Stream);
  if (m_Module) (void)hipModuleUnload(m_Module);
}

int Particles::Initialize(const char* path) {
  if (m_Module != nullptr) return fx::fail(FX_ESTATE, "Initialize: already initialized");
  FX_HIP(hipModuleLoad(&m_Module, path), "Initialize");
  m_CodePath = path;
  for (int f = 0; f < FUNC_MAX; f++)
    FX_HIP(hipModuleGetFunction(&m_Func[f], m_Module, kKernelNames[f]), kKernelNames[f]);
  FX_TRY(m_Points.AssignToGPU("FPnts", m_Module));
  FX_TRY(m_PointsTemp.AssignToGPU("FPntTmp", m_Module));
  FX_TRY(m_Accel.AssignToGPU("FAccel", m_Module));
  size_t len = 0;
  FX_HIP(hipModuleGetGlobal(&m_cuParams, &len, m_Module, "FParams"), "Initialize");
  if (len != sizeof(FParams_t))
    return fx::fail(FX_EMISMATCH, "Initialize: FParams has %zu bytes, FParams_t has %zu", len,
                    sizeof(FParams_t));
  FX_HIP(hipModuleGetGlobal(reinterpret_cast<hipDeviceptr_t*>(&m_NbrWordsSym), &len, m_Module, "FNbrWords"),
         "Initialize");
  FX_HIP(hipModuleGetGlobal(reinterpret_cast<hipDeviceptr_t*>(&m_SortSrcSym), &len, m_Module, "FSortSrc"),
         "Initialize");
  if (len != sizeof(void*)) return fx::fail(FX_EMISMATCH, "Initialize: FSortSrc has %zu bytes", len);
  // this host allocates the tail slot and runs the index sort (a host that
  // binds the module as the reference does leaves FHostCaps at 0)
  hipDeviceptr_t caps = nullptr;
  FX_HIP(hipModuleGetGlobal(&caps, &len, m_Module, "FHostCaps"), "Initialize");
  if (len != sizeof(int)) return fx::fail(FX_EMISMATCH, "Initialize: FHostCaps has %zu bytes", len);
  int capv = FX_CAP_TAIL | FX_CAP_RANKED;
  FX_HIP(hipMemcpyHtoD(caps, &capv, sizeof(int)), "Initialize");
  FX_HIP(hipStreamCreate(&m_Stream), "Initialize");
  FX_HIP(hipMalloc(&m_Tables, 2 * sizeof(cuDataX)), "Initialize");
  FX_HIP(hipMalloc((void**)&m_Clock, 2 * sizeof(float)), "Initialize");
  FX_HIP(hipMalloc(&m_Stats, 2 * sizeof(unsigned long long)), "Initialize");
  FX_HIP(hipMalloc((void**)&m_WorkQ, 8 * FX_NUM_STAGES * sizeof(uint32_t)), "Initialize");
  FX_HIP(hipModuleGetGlobal(reinterpret_cast<hipDeviceptr_t*>(&m_PptSym), &len, m_Module, "FPpt"), "Initialize");
  if (len != 8 * sizeof(int)) return fx::fail(FX_EMISMATCH, "Initialize: FPpt has %zu bytes", len);
  {
    int dev = 0;
    FX_HIP(hipGetDevice(&dev), "Initialize");
    FX_HIP(hipDeviceGetAttribute(&m_NumCU, hipDeviceAttributeMultiprocessorCount, dev), "Initialize");
  }
  FX_TRY(UploadPpt());
  FX_TRY(fx::scratch_setup(&m_ScanScratch));
  m_Params.numThreads = 256;
  return FX_OK;
}

int Particles::Launch(int fid, unsigned blocks, unsigned threads, void** args) {
  if (blocks == 0) return FX_OK;
  FX_HIP(hipModuleLaunchKernel(m_Func[fid], blocks, 1, 1, threads, 1, 1, 0, m_Stream, args, nullptr),
         kKernelNames[fid]);
  return FX_OK;
}

int Particles::LaunchP(int fid, int stage, void** args) {
  const int waves = (mNumPoints + 63) / 64 * 64;
  const int threads = m_StageThreads[stage] < waves ? m_StageThreads[stage] : waves;
  if (threads <= 0) return FX_OK;
  const int per_block = threads * m_StagePpt[stage];
  return Launch(fid, (unsigned)((mNumPoints + per_block - 1) / per_block), (unsigned)threads, args);
}

int Particles::UploadPpt() {
  int v[8] = {1, 1, 1, 1, 1, 1, 1, 1};
  for (int k = 0; k < FX_NUM_STAGES; k++) v[k] = m_StagePpt[k];
  FX_HIP(hipStreamSynchronize(m_Stream ? m_Stream : nullptr), "ppt");
  FX_HIP(hipMemcpyHtoD(reinterpret_cast<hipDeviceptr_t>(m_PptSym), v, sizeof(v)), "ppt");
  return FX_OK;
}

int Particles::SetStagePpt(int stage, int ppt) {
  if (stage < 0 || stage >= FX_NUM_STAGES) return fx::fail(FX_EINVAL, "SetStagePpt: bad stage %d", stage);
  if (ppt < 1 || ppt > 8) return fx::fail(FX_EINVAL, "SetStagePpt: %d particles per thread is not in [1, 8]", ppt);
  if (ppt != 1 && stage != FX_STAGE_INSERT && stage != FX_STAGE_SORT && stage != FX_STAGE_ADVANCE)
    return fx::fail(FX_EINVAL, "SetStagePpt: stage %d maps one particle to one lane (the pair stages batch work "
                    "per block and per wave instead)", stage);
  if (m_StagePpt[stage] == ppt) return FX_OK;
  ReleaseGraph();
  m_StagePpt[stage] = ppt;
  FX_TRY(UploadPpt());
  return BuildBlockTables();
}

int Particles::SetPersistent(bool on) {
  if (on == m_Persistent) return FX_OK;
  ReleaseGraph();
  m_Persistent = on;
  return FX_OK;
}

int Particles::ResidentBlocks(int fid, int threads, int* blocks) {
  const auto key = std::make_pair(fid, threads);
  auto it = m_Resident.find(key);
  if (it == m_Resident.end()) {
    int per_cu = 0;
    FX_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, m_Func[fid], threads, 0), "occupancy");
    if (per_cu < 1) per_cu = 1;
    it = m_Resident.emplace(key, per_cu * (m_NumCU > 0 ? m_NumCU : 1)).first;
  }
  *blocks = it->second;
  return FX_OK;
}

int Particles::Restart(const fx_scene& sc) {
  if (m_Module == nullptr) return fx::fail(FX_ESTATE, "Restart: call Initialize first");
  FX_HIP(hipStreamSynchronize(m_Stream), "Restart");
  ReleaseGraph();
  AdjRelease();
  free_batch(m_DParams, m_DPBase, m_DCBase, m_BTab);
  m_Batch = false;
  m_SceneParams.clear();
  const int threads = m_Params.numThreads > 0 ? m_Params.numThreads : 256;
  FParams_t P;
  FX_TRY(fx::scene_derive(sc, P));
  m_Params = P;
  mNumPoints = 0;
  mMaxPoints = sc.num_points;
  m_ScanCells = m_Params.gridTotal + 1;
  m_PBase = {0, sc.num_points};
  m_CBase = {0, m_ScanCells};
  FX_TRY(ReallocateParticles(mMaxPoints));
  fx::set_geometry(m_Params, threads);
  FX_TRY(RebuildAccelGrid());

  // AddPointsInVolume: lattice into the host copies, other fields zero
  const int n = fx::scene_fill(sc, reinterpret_cast<float*>(m_Points.bufF3(FPOS)), m_Points.bufUI(FCLR),
                               mMaxPoints);
  if (n < 0) return n;
  if (n != mMaxPoints)
    return fx::fail(FX_ERANGE, "Restart: lattice holds %d of %d particles", n, mMaxPoints);
  return FinishRestart(n);
}

int Particles::RestartBatch(const fx_scene* scenes, int count) {
  if (m_Module == nullptr) return fx::fail(FX_ESTATE, "RestartBatch: call Initialize first");
  if (scenes == nullptr || count <= 0) return fx::fail(FX_EINVAL, "RestartBatch: no scenes");
  FX_HIP(hipStreamSynchronize(m_Stream), "RestartBatch");
  ReleaseGraph();
  AdjRelease();
  free_batch(m_DParams, m_DPBase, m_DCBase, m_BTab);
  std::vector<FParams_t> ps(count);
  std::vector<int> pb(count + 1, 0), cb(count + 1, 0);
  for (int s = 0; s < count; s++) {
    FX_TRY(fx::scene_derive(scenes[s], ps[s]));
    if ((int64_t)pb[s] + scenes[s].num_points > INT32_MAX ||
        (int64_t)cb[s] + ps[s].gridTotal + 1 > (int64_t)1024 * 1024 * 1024 - 1)
      return fx::fail(FX_ERANGE, "RestartBatch: the batch exceeds 2^31 particles or 1024^3 cells");
    pb[s + 1] = pb[s] + scenes[s].num_points;
    cb[s + 1] = cb[s] + ps[s].gridTotal + 1;
  }
  const int threads = m_Params.numThreads > 0 ? m_Params.numThreads : 256;
  m_Batch = true;
  m_SceneParams = ps;
  m_PBase = pb;
  m_CBase = cb;
  m_Params = ps[0];  // time, dt and the single-scene symbol
  mNumPoints = 0;
  mMaxPoints = pb[count];
  m_ScanCells = cb[count];
  m_Params.pnum = mMaxPoints;
  FX_TRY(ReallocateParticles(mMaxPoints));
  fx::set_geometry(m_Params, threads);
  FX_TRY(RebuildAccelGrid());
  for (int s = 0; s < count; s++) {
    ps[s].pnum = scenes[s].num_points;
    const int n = fx::scene_fill(scenes[s], reinterpret_cast<float*>(m_Points.bufF3(FPOS, pb[s])),
                                 m_Points.bufUI(FCLR, pb[s]), scenes[s].num_points);
    if (n < 0) return n;
    if (n != scenes[s].num_points)
      return fx::fail(FX_ERANGE, "RestartBatch: scene %d lattice holds %d of %d particles", s, n,
                      scenes[s].num_points);
  }
  m_SceneParams = ps;
  FX_HIP(hipMalloc(&m_DParams, sizeof(FParams_t) * count), "RestartBatch");
  FX_HIP(hipMalloc(&m_DPBase, sizeof(int) * (count + 1)), "RestartBatch");
  FX_HIP(hipMalloc(&m_DCBase, sizeof(int) * (count + 1)), "RestartBatch");
  FX_HIP(hipMemcpyHtoD(m_DParams, ps.data(), sizeof(FParams_t) * count), "RestartBatch");
  FX_HIP(hipMemcpyHtoD(m_DPBase, pb.data(), sizeof(int) * (count + 1)), "RestartBatch");
  FX_HIP(hipMemcpyHtoD(m_DCBase, cb.data(), sizeof(int) * (count + 1)), "RestartBatch");
  FX_TRY(BuildBlockTables());
  return FinishRestart(mMaxPoints);
}

int Particles::BuildBlockTables() {
  // (scene, block inside the scene) for every block of every stage's launch
  for (auto& e : m_BTab)
    if (e.second.first) (void)hipFree(e.second.first);
  m_BTab.clear();
  if (!m_Batch) return FX_OK;
  for (int k = 0; k < FX_NUM_STAGES; k++) {
    const int T = m_StageThreads[k] * m_StagePpt[k];  // particles per block
    if (m_BTab.count(T)) continue;
    std::vector<int> tab;
    for (int s = 0; s + 1 < (int)m_PBase.size(); s++) {
      const int nb = (m_PBase[s + 1] - m_PBase[s] + T - 1) / T;
      for (int b = 0; b < nb; b++) {
        tab.push_back(s);
        tab.push_back(b);
      }
    }
    void* d = nullptr;
    FX_HIP(hipMalloc(&d, tab.size() * sizeof(int) + 8), "BuildBlockTables");
    FX_HIP(hipMemcpyHtoD(d, tab.data(), tab.size() * sizeof(int)), "BuildBlockTables");
    m_BTab[T] = {d, (int)(tab.size() / 2)};
  }
  return FX_OK;
}

int Particles::BatchLaunch(int fid, int fid_p, int stage, void* extra0, void* extra1, void* extra2) {
  // One launch over every scene of the batch: one block per block-table
  // entry, or (persistent) as many blocks as the device holds at once,
  // capped by the entries, pulling entries from the stage's claim counters,
  // which a fill kernel clears first.
  const int T = m_StageThreads[stage];
  const auto it = m_BTab.find(T * m_StagePpt[stage]);
  if (it == m_BTab.end()) return fx::fail(FX_ESTATE, "batch launch: no block table for stage %d", stage);
  FxBatchArgs B{m_DParams, m_DPBase, m_DCBase, it->second.first, nullptr, it->second.second, 0};
  unsigned grid = (unsigned)it->second.second;
  if (fid_p > 0 && m_Persistent) {
    fid = fid_p;
    B.queue = m_WorkQ + 8 * stage;
    int resident = 0;
    FX_TRY(ResidentBlocks(fid, T, &resident));
    if ((unsigned)resident < grid) grid = (unsigned)resident;
    int eight = 8;
    unsigned zero = 0;
    void* q = B.queue;
    void* fargs[] = {&q, &zero, &eight};
    FX_TRY(Launch(FUNC_FILL, 1, 64, fargs));
  }
  void* args[4] = {&B, extra0, extra1, extra2};
  return Launch(fid, grid, (unsigned)T, args);
}

int Particles::SceneRange(int s, int* p0, int* np, int* c0, int* nc) const {
  if (s < 0 || s + 1 >= (int)m_PBase.size()) return fx::fail(FX_EINVAL, "scene %d out of range", s);
  *p0 = m_PBase[s];
  *np = m_PBase[s + 1] - m_PBase[s];
  *c0 = m_CBase[s];
  *nc = m_CBase[s + 1] - m_CBase[s];
  return FX_OK;
}

int Particles::SceneParams(int s, FParams_t* out) const {
  if (s < 0 || s >= NumScenes()) return fx::fail(FX_EINVAL, "scene %d out of range", s);
  *out = m_Batch ? m_SceneParams[s] : m_Params;
  if (m_Batch) out->time = m_Params.time;
  return FX_OK;
}

int Particles::FinishRestart(int n) {
  memset(m_Points.bufF3(FVEL), 0, sizeof(fx_float3) * (size_t)n);
  memset(m_Points.bufF3(FVEVAL), 0, sizeof(fx_float3) * (size_t)n);
  memset(m_Points.bufF3(FFORCE), 0, sizeof(fx_float3) * (size_t)n);
  memset(m_Points.bufF(FPRESS), 0, sizeof(float) * (size_t)n);
  memset(m_Points.bufI(FGCELL), 0, sizeof(int) * (size_t)n);
  memset(m_Points.bufI(FGNDX), 0, sizeof(int) * (size_t)n);
  for (int i = 0; i < n; i++) m_Points.bufUI(FID)[i] = (unsigned)i;
  mNumPoints = n;
  m_Points.SetNum(n);
  m_PointsTemp.SetNum(n);

  FX_TRY(UpdateParams());
  FX_TRY(CommitAll());
  m_Frame = 0;
  m_Done = -1;
  return FX_OK;
}

int Particles::ReallocateParticles(int maxpnt) {
  // particles.cpp:267-300, plus the FID slot; GL interop flags become device memory
  if (maxpnt < 0 || maxpnt >= (1 << FX_MEMBER_BITS))
    return fx::fail(FX_ERANGE, "ReallocateParticles: %d particles (the ranked sort's member entries hold "
                               "fewer than 2^%d)", maxpnt, FX_MEMBER_BITS);
  mMaxPoints = maxpnt;
  m_Params.pnum = maxpnt;
  m_Points.DeleteAllBuffers();
  const uchar fl = DT_CPU | DT_CUMEM;
  FX_TRY(m_Points.AddBuffer(FPOS, "pos", sizeof(fx_float3), maxpnt, fl));
  FX_TRY(m_Points.AddBuffer(FCLR, "clr", sizeof(unsigned), maxpnt, fl));
  FX_TRY(m_Points.AddBuffer(FVEL, "vel", sizeof(fx_float3), maxpnt, fl));
  FX_TRY(m_Points.AddBuffer(FVEVAL, "veval", sizeof(fx_float3), maxpnt, fl));
  FX_TRY(m_Points.AddBuffer(FPRESS, "press", sizeof(float), maxpnt, fl));
  FX_TRY(m_Points.AddBuffer(FFORCE, "force", sizeof(fx_float3), maxpnt, fl));
  FX_TRY(m_Points.AddBuffer(FGCELL, "gcell", sizeof(int), maxpnt, fl));
  FX_TRY(m_Points.AddBuffer(FGNDX, "gndx", sizeof(int), maxpnt, fl));
  FX_TRY(m_Points.AddBuffer(FID, "id", sizeof(unsigned), maxpnt, fl));
  m_Points.SetBufferUsage(FPOS, DT_FLOAT3);
  m_Points.SetBufferUsage(FCLR, DT_UINT);
  m_Points.SetBufferUsage(FVEL, DT_FLOAT3);
  m_Points.SetBufferUsage(FVEVAL, DT_FLOAT3);
  m_Points.SetBufferUsage(FPRESS, DT_FLOAT);
  m_Points.SetBufferUsage(FFORCE, DT_FLOAT3);
  m_Points.SetBufferUsage(FGCELL, DT_INT);
  m_Points.SetBufferUsage(FGNDX, DT_INT);
  m_Points.SetBufferUsage(FID, DT_UINT);
  FX_TRY(m_PointsTemp.MatchAllBuffers(&m_Points, DT_CUMEM));
  m_Parity = 0;
  return FX_OK;
}

int Particles::RebuildAccelGrid() {
  // particles.cpp:548-626; the counts and offsets carry one extra slot
  // (index gridTotal) for out-of-range particles.
  FX_TRY(fx::rebuild_grid(m_Params));
  const int numElem1 = m_ScanCells;  // every scene's cells and tail bucket
  const int numElem2 = numElem1 / SCAN_ELEMS + 1;
  const int numElem3 = numElem2 / SCAN_ELEMS + 1;
  if (numElem3 > SCAN_ELEMS) return fx::fail(FX_ERANGE, "RebuildAccelGrid: %d cells exceed the scan", numElem1);
  m_Accel.DeleteAllBuffers();
  FX_TRY(m_Accel.AddBuffer(AGRID, "grid", sizeof(unsigned), mMaxPoints, DT_CUMEM));
  FX_TRY(m_Accel.AddBuffer(AGRIDCNT, "gridcnt", sizeof(unsigned), numElem1, DT_CUMEM));
  FX_TRY(m_Accel.AddBuffer(AGRIDOFF, "gridoff", sizeof(unsigned), numElem1, DT_CUMEM));
  FX_TRY(m_Accel.AddBuffer(AAUXARRAY1, "aux1", sizeof(unsigned), numElem2, DT_CUMEM));
  FX_TRY(m_Accel.AddBuffer(AAUXSCAN1, "scan1", sizeof(unsigned), numElem2, DT_CUMEM));
  FX_TRY(m_Accel.AddBuffer(AAUXARRAY2, "aux2", sizeof(unsigned), numElem3, DT_CUMEM));
  FX_TRY(m_Accel.AddBuffer(AAUXSCAN2, "scan2", sizeof(unsigned), numElem3, DT_CUMEM));
  FX_TRY(m_Accel.AddBuffer(ANBRBITS, "nbrbits", sizeof(unsigned), (uint64_t)FX_NBR_WORDS * mMaxPoints, DT_CUMEM));
  FX_TRY(m_Accel.AddBuffer(AFREC, "frec", 32, mMaxPoints, DT_CUMEM));
  FX_TRY(m_Accel.AddBuffer(ASLABSTART, "slabstart", sizeof(unsigned), (uint64_t)FX_SLABS * numElem1, DT_CUMEM));
  // out-of-range flags of the particles and their scan levels
  const int tail1 = mMaxPoints > 0 ? mMaxPoints : 1;
  const int tail2 = tail1 / SCAN_ELEMS + 1;
  const int tail3 = tail2 / SCAN_ELEMS + 1;
  FX_TRY(m_Accel.AddBuffer(ATAILFLAG, "tailflag", sizeof(unsigned), tail1, DT_CUMEM));
  FX_TRY(m_Accel.AddBuffer(ATAILSCAN, "tailscan", sizeof(unsigned), tail1, DT_CUMEM));
  FX_TRY(m_Accel.AddBuffer(ATAILAUX1, "tailaux1", sizeof(unsigned), tail2, DT_CUMEM));
  FX_TRY(m_Accel.AddBuffer(ATAILAUXS1, "tailscan1", sizeof(unsigned), tail2, DT_CUMEM));
  FX_TRY(m_Accel.AddBuffer(ATAILAUX2, "tailaux2", sizeof(unsigned), tail3, DT_CUMEM));
  FX_TRY(m_Accel.AddBuffer(ATAILAUXS2, "tailscan2", sizeof(unsigned), tail3, DT_CUMEM));
  for (int b = 0; b <= AAUXSCAN2; b++) m_Accel.SetBufferUsage(b, DT_UINT);
  for (int b = ATAILFLAG; b <= ATAILAUXS2; b++) m_Accel.SetBufferUsage(b, DT_UINT);
  FX_TRY(UpdateGPUAccess());
  return FX_OK;
}

int Particles::UpdateGPUAccess() {
  // particles.cpp:649-655, plus the device copy of both sets' tables used by
  // the per-step exchange
  cuDataX t[2];
  m_Points.FillTable(&t[m_Parity]);
  m_PointsTemp.FillTable(&t[m_Parity ^ 1]);
  FX_HIP(hipMemcpyHtoD(m_Tables, t, sizeof(t)), "UpdateGPUAccess");
  FX_TRY(m_Points.UpdateGPUAccess());
  FX_TRY(m_PointsTemp.UpdateGPUAccess());
  FX_TRY(m_Accel.UpdateGPUAccess());
  return FX_OK;
}

int Particles::UpdateParams() {
  fx::update_params(m_Params);
  FX_HIP(hipStreamSynchronize(m_Stream), "UpdateParams");
  FX_HIP(hipMemcpyHtoD(m_cuParams, &m_Params, sizeof(FParams_t)), "UpdateParams");
  return FX_OK;
}

int Particles::SetParams(const FParams_t& p) {
  if (mMaxPoints == 0) return fx::fail(FX_ESTATE, "SetParams: call Restart first");
  const FParams_t& q = m_Params;
  if (p.gridRes.x != q.gridRes.x || p.gridRes.y != q.gridRes.y || p.gridRes.z != q.gridRes.z ||
      memcmp(&p.gridMin, &q.gridMin, sizeof(fx_float3)) != 0 ||
      memcmp(&p.gridDelta, &q.gridDelta, sizeof(fx_float3)) != 0 || p.gridSrch != q.gridSrch ||
      p.pnum != q.pnum || p.psmoothradius != q.psmoothradius || p.sim_scale != q.sim_scale)
    return fx::fail(FX_ESTATE, "SetParams: the grid, smoothing radius, scale and count are fixed by Restart");
  if (m_Batch) return fx::fail(FX_ESTATE, "SetParams: a batch takes its parameters from its scenes");
  ReleaseGraph();
  const int threads = m_Params.numThreads;
  m_Params = p;
  fx::set_geometry(m_Params, threads);
  return UpdateParams();
}

int Particles::SetBlockSize(int threads) {
  if (threads < 64 || threads > 1024 || threads % 64 != 0)
    return fx::fail(FX_EINVAL, "SetBlockSize: %d is not a multiple of 64 in [64, 1024]", threads);
  ReleaseGraph();
  m_Params.numThreads = threads;
  for (int k = 0; k < FX_NUM_STAGES; k++) m_StageThreads[k] = threads;
  if (mMaxPoints > 0) fx::set_geometry(m_Params, threads);
  return BuildBlockTables();
}

int Particles::SetStageBlockSize(int stage, int threads) {
  if (stage < 0 || stage >= FX_NUM_STAGES) return fx::fail(FX_EINVAL, "SetStageBlockSize: bad stage %d", stage);
  if (threads < 64 || threads > 1024 || threads % 64 != 0)
    return fx::fail(FX_EINVAL, "SetStageBlockSize: %d is not a multiple of 64 in [64, 1024]", threads);
  ReleaseGraph();
  m_StageThreads[stage] = threads;
  return BuildBlockTables();
}

int Particles::CommitAll() {
  FX_HIP(hipStreamSynchronize(m_Stream), "CommitAll");
  return m_Points.CommitAll(m_Stream);
}

// ------------------------------------------------------------------ stages

int Particles::InsertParticles() {
  // particles.cpp:716-730.  Only the counts need clearing: gcell and gndx of
  // every particle are written by the kernel.
  int ncell = m_ScanCells;
  unsigned zero = 0;
  void* cnt = m_Accel.gpu(AGRIDCNT);
  void* fargs[] = {&cnt, &zero, &ncell};
  FX_TRY(Launch(FUNC_FILL, (unsigned)((ncell + 255) / 256), 256, fargs));
  if (m_Batch) {
    FX_TRY(BatchLaunch(FUNC_INSERT_B, FUNC_INSERT_BP, FX_STAGE_INSERT, nullptr, nullptr, nullptr));
  } else {
    void* args[] = {&mNumPoints};
    FX_TRY(LaunchP(FUNC_INSERT, FX_STAGE_INSERT, args));
  }
  m_Done = FX_STAGE_INSERT;
  return FX_OK;
}

int Particles::PrefixScanParticles() {
  // particles.cpp:774-815, over gridTotal+1 counts; then the out-of-range
  // flags, whose exclusive scan ranks the tail bucket in particle order.
  // Both scans in two launches (block totals first, into the first
  // auxiliary level of each, which holds one entry per 1024 elements); the
  // reference's three-level prefixSum/prefixFixup for a scan beyond
  // FX_SCAN2_MAX_BLOCKS blocks.
  FX_TRY(ScanTwoLaunch(m_Accel.gpu(AGRIDCNT), m_Accel.gpu(AGRIDOFF), m_Accel.gpu(AAUXARRAY1), m_ScanCells,
                       m_Accel.gpu(ATAILFLAG), m_Accel.gpu(ATAILSCAN), m_Accel.gpu(ATAILAUX1), mNumPoints));
  m_Done = FX_STAGE_PREFIX;
  return FX_OK;
}

namespace {
// kernel argument of fxScanSums / fxScanApply (same layout as the kernel's)
struct ScanJob {
  const void* in;
  void* out;
  void* sums;
  int n;
  int nb;
};
struct ScanJobs {
  ScanJob j[2];
};
}  // namespace

int Particles::ScanTwoLaunch(void* in, void* out, void* sums, int n, void* in2, void* out2, void* sums2, int n2) {
  const int nb = (n + FX_SCAN2_BLOCK - 1) / FX_SCAN2_BLOCK;
  const int nb2 = (n2 + FX_SCAN2_BLOCK - 1) / FX_SCAN2_BLOCK;
  if (nb > FX_SCAN2_MAX_BLOCKS || nb2 > FX_SCAN2_MAX_BLOCKS) {
    // each block of the second launch reads the totals before it: beyond a
    // few thousand blocks the three-level scan reads less
    auto one = [&](void* a, void* o, void* s, int m) -> int {
      if ((m + FX_SCAN2_BLOCK - 1) / FX_SCAN2_BLOCK <= FX_SCAN2_MAX_BLOCKS) return ScanTwoLaunch(a, o, s, m);
      if (a == m_Accel.gpu(AGRIDCNT))
        return ScanLevels(a, o, m_Accel.gpu(AAUXARRAY1), m_Accel.gpu(AAUXSCAN1), m_Accel.gpu(AAUXARRAY2),
                          m_Accel.gpu(AAUXSCAN2), m);
      return ScanLevels(a, o, m_Accel.gpu(ATAILAUX1), m_Accel.gpu(ATAILAUXS1), m_Accel.gpu(ATAILAUX2),
                        m_Accel.gpu(ATAILAUXS2), m);
    };
    FX_TRY(one(in, out, sums, n));
    return n2 > 0 ? one(in2, out2, sums2, n2) : FX_OK;
  }
  ScanJobs J = {{{in, out, sums, n, nb}, {in2, out2, sums2, n2, nb2}}};
  if (nb + nb2 == 0) return FX_OK;
  void* a[] = {&J};
  FX_TRY(Launch(FUNC_SCAN_SUMS, (unsigned)(nb + nb2), FX_SCAN2_THREADS, a));
  return Launch(FUNC_SCAN_APPLY, (unsigned)(nb + nb2), FX_SCAN2_THREADS, a);
}

int Particles::ScanLevels(void* array1, void* scan1, void* array2, void* scan2, void* array3, void* scan3,
                          int n) {
  // exclusive scan in up to three levels of 1024-element blocks
  int numElem1 = n;
  int numElem2 = numElem1 / SCAN_ELEMS + 1;
  int numElem3 = numElem2 / SCAN_ELEMS + 1;
  if (numElem3 > SCAN_ELEMS) return fx::fail(FX_ERANGE, "scan: %d elements exceed 1024^3", n);
  int zon = 1;
  void* argsA[] = {&array1, &scan1, &array2, &numElem1, &zon};
  FX_TRY(Launch(FUNC_FPREFIXSUM, (unsigned)numElem2, SCAN_BLOCKSIZE, argsA));
  void* argsB[] = {&array2, &scan2, &array3, &numElem2, &zon};
  FX_TRY(Launch(FUNC_FPREFIXSUM, (unsigned)numElem3, SCAN_BLOCKSIZE, argsB));
  if (numElem3 > 1) {
    void* nptr = nullptr;
    void* argsC[] = {&array3, &scan3, &nptr, &numElem3, &zon};
    FX_TRY(Launch(FUNC_FPREFIXSUM, 1, SCAN_BLOCKSIZE, argsC));
    void* argsD[] = {&scan2, &scan3, &numElem2};
    FX_TRY(Launch(FUNC_FPREFIXFIXUP, (unsigned)numElem3, SCAN_BLOCKSIZE, argsD));
  }
  void* argsE[] = {&scan1, &scan2, &numElem1};
  return Launch(FUNC_FPREFIXFIXUP, (unsigned)numElem2, SCAN_BLOCKSIZE, argsE);
}

int Particles::PublishTables() {
  // FPnts <- table of the set m_Points now holds, FPntTmp <- the other one;
  // in a captured step the same launch moves the step clock (the advance of
  // this step reads clock[0])
  void* bank = m_Tables;
  void* pnts = m_Points.SymbolAddress();
  void* tmp = m_PointsTemp.SymbolAddress();
  void* clk = m_Capturing ? (void*)m_Clock : nullptr;
  float dt = m_Params.dt;
  int parity = m_Parity;
  void* args[] = {&bank, &parity, &pnts, &tmp, &clk, &dt};
  return Launch(FUNC_PUBLISH, 1, 64, args);
}

int Particles::SetDeviceClock() {
  // clock[1] is the time of the next step a captured graph runs
  FX_HIP(hipMemcpyHtoD(m_Clock + 1, &m_Params.time, sizeof(float)), "device clock");
  return FX_OK;
}

int Particles::CountingSort() {
  // particles.cpp:832-842.  CopyAllToTemp is replaced by exchanging the two
  // sets: FPntTmp then holds the pre-sort data, FPnts receives the sort.
  FX_TRY(m_Points.SwapGPU(&m_PointsTemp));
  m_Parity ^= 1;
  FX_TRY(PublishTables());
  if (m_Batch) {
    FX_TRY(BatchLaunch(FUNC_SORT_INDEX_B, FUNC_SORT_INDEX_BP, FX_STAGE_SORT, nullptr, nullptr, nullptr));
    FX_TRY(BatchLaunch(FUNC_SORT_FULL_B, FUNC_SORT_FULL_BP, FX_STAGE_SORT, nullptr, nullptr, nullptr));
  } else {
    void* args[] = {&mNumPoints};
    FX_TRY(LaunchP(FUNC_SORT_INDEX, FX_STAGE_SORT, args));
    FX_TRY(LaunchP(FUNC_COUNTING_SORT, FX_STAGE_SORT, args));
  }
  m_Done = FX_STAGE_SORT;
  return FX_OK;
}

int Particles::ComputePressure() {
  if (m_Batch) {
    FX_TRY(BatchLaunch(FUNC_PRESS_B, FUNC_PRESS_BP, FX_STAGE_PRESSURE, &mNumPoints, nullptr, nullptr));
  } else {
    void* args[] = {&mNumPoints};
    FX_TRY(LaunchP(FUNC_COMPUTE_PRESS, FX_STAGE_PRESSURE, args));
  }
  m_Done = FX_STAGE_PRESSURE;
  return FX_OK;
}

int Particles::ComputeForce() {
  if (m_Batch) {
    FX_TRY(BatchLaunch(FUNC_FORCE_B, FUNC_FORCE_BP, FX_STAGE_FORCE, &mNumPoints, nullptr, nullptr));
  } else {
    void* args[] = {&mNumPoints};
    FX_TRY(LaunchP(FUNC_COMPUTE_FORCE, FX_STAGE_FORCE, args));
  }
  m_Done = FX_STAGE_FORCE;
  return FX_OK;
}

int Particles::Advance() {
  // particles.cpp:413-418 (launched over pnum, which equals mNumPoints here)
  float dt = m_Params.dt, ss = m_Params.sim_scale;
  int n = m_Params.pnum;
  // a captured step reads its time from the device clock, which the sort
  // stage's table exchange (PublishTables) has set
  if (m_Batch) {
    float time = m_Params.time;
    void* clk = m_Capturing ? (void*)m_Clock : nullptr;
    return BatchLaunch(FUNC_ADVANCE_B, FUNC_ADVANCE_BP, FX_STAGE_ADVANCE, &clk, &time, nullptr);
  }
  if (m_Capturing) {
    void* args[] = {&m_Clock, &dt, &ss, &n};
    return LaunchP(FUNC_ADVANCE_CLOCK, FX_STAGE_ADVANCE, args);
  }
  float time = m_Params.time;
  void* args[] = {&time, &dt, &ss, &n};
  return LaunchP(FUNC_ADVANCE, FX_STAGE_ADVANCE, args);
}

void Particles::AdvanceTime() {
  m_Params.time += m_Params.dt;
  m_Frame++;
}

int Particles::Stage(int stage) {
  if (mNumPoints == 0) return fx::fail(FX_ESTATE, "stage: call Restart first");
  // Each stage indexes with tables that the stages before it in the same step
  // write (counts, offsets, sorted cells, the neighbour bitmask); a call out of
  // that order is refused rather than run on stale tables.  Insert and advance
  // read only per-particle data.
  const bool ok = stage == FX_STAGE_PREFIX ? (m_Done == FX_STAGE_INSERT || m_Done == FX_STAGE_PREFIX)
                  : stage == FX_STAGE_SORT ? m_Done == FX_STAGE_PREFIX
                  : stage == FX_STAGE_PRESSURE ? m_Done >= FX_STAGE_SORT
                  : stage == FX_STAGE_FORCE ? m_Done >= FX_STAGE_PRESSURE
                                            : true;
  if (!ok)
    return fx::fail(FX_ESTATE, "stage %d: out of order (last stage run since the insert: %d); a step runs insert, "
                    "prefix, sort, pressure, force, advance", stage, m_Done);
  switch (stage) {
    case FX_STAGE_INSERT: return InsertParticles();
    case FX_STAGE_PREFIX: return PrefixScanParticles();
    case FX_STAGE_SORT: return CountingSort();
    case FX_STAGE_PRESSURE: return ComputePressure();
    case FX_STAGE_FORCE: return ComputeForce();
    case FX_STAGE_ADVANCE: {
      FX_TRY(Advance());
      AdvanceTime();
      m_Done = FX_STAGE_ADVANCE;
      return FX_OK;
    }
    default: return fx::fail(FX_EINVAL, "stage: unknown stage %d", stage);
  }
}

int Particles::Run() {
  if (mNumPoints == 0) return fx::fail(FX_ESTATE, "Run: call Restart first");
  FX_TRY(InsertParticles());
  FX_TRY(PrefixScanParticles());
  FX_TRY(CountingSort());
  FX_TRY(ComputePressure());
  FX_TRY(ComputeForce());
  FX_TRY(Advance());
  AdvanceTime();
  m_Done = FX_STAGE_ADVANCE;
  return FX_OK;
}

// ------------------------------------------------------------------ graphs

void Particles::ReleaseGraph() {
  if (m_GraphExec) (void)hipGraphExecDestroy(m_GraphExec);
  if (m_Graph) (void)hipGraphDestroy(m_Graph);
  m_GraphExec = nullptr;
  m_Graph = nullptr;
}

int Particles::CaptureGraph() {
  if (mNumPoints == 0) return fx::fail(FX_ESTATE, "CaptureGraph: call Restart first");
  FX_HIP(hipStreamSynchronize(m_Stream), "CaptureGraph");
  ReleaseGraph();
  const int parity0 = m_Parity, done0 = m_Done;
  FX_HIP(hipStreamBeginCapture(m_Stream, hipStreamCaptureModeThreadLocal), "CaptureGraph");
  m_Capturing = true;
  int st = FX_OK;
  for (int s = 0; s < 2 && st == FX_OK; s++) {
    st = InsertParticles();
    if (st == FX_OK) st = PrefixScanParticles();
    if (st == FX_OK) st = CountingSort();
    if (st == FX_OK) st = ComputePressure();
    if (st == FX_OK) st = ComputeForce();
    if (st == FX_OK) st = Advance();
  }
  m_Capturing = false;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(m_Stream, &g);
  if (st != FX_OK || e != hipSuccess) {
    // nothing of a failed capture ran: restore the host view of the sets
    if (g) (void)hipGraphDestroy(g);
    if (m_Parity != parity0) {
      (void)m_Points.SwapGPU(&m_PointsTemp);
      m_Parity = parity0;
    }
    m_Done = done0;
    return st != FX_OK ? st : fx::hip_fail(e, "CaptureGraph", "hipStreamEndCapture");
  }
  m_Done = done0;  // recorded, not run
  m_Graph = g;
  m_GraphParity = m_Parity;
  FX_HIP(hipGraphInstantiate(&m_GraphExec, m_Graph, nullptr, nullptr, 0), "CaptureGraph");
  return FX_OK;
}

int Particles::LaunchGraph(int nlaunch) {
  // the graph replays the set exchange it was captured with
  if (m_GraphExec == nullptr || m_GraphParity != m_Parity) FX_TRY(CaptureGraph());
  // the device clock starts from the host time of the first replayed step
  FX_HIP(hipStreamSynchronize(m_Stream), "LaunchGraph");
  FX_TRY(SetDeviceClock());
  for (int k = 0; k < nlaunch; k++) {
    FX_HIP(hipGraphLaunch(m_GraphExec, m_Stream), "LaunchGraph");
    AdvanceTime();
    AdvanceTime();
    m_Done = FX_STAGE_ADVANCE;
  }
  return FX_OK;
}

// ---------------------------------------------------------------- transfers

int Particles::CheckIds(const uint32_t* id, int64_t n, const char* who) {
  std::vector<uint8_t> seen((size_t)n, 0);
  for (int64_t k = 0; k < n; k++) {
    if (id[k] >= (uint64_t)n || seen[id[k]])
      return fx::fail(FX_EINVAL, "%s: ids are not a permutation of 0..%lld (entry %lld = %u)", who, (long long)n - 1,
                      (long long)k, id[k]);
    seen[id[k]] = 1;
  }
  return FX_OK;
}

int Particles::Upload(int buf, const void* host, int64_t count) {
  DataPtr* b = m_Points.GetBuffer(buf);
  if (b == nullptr || host == nullptr) return fx::fail(FX_EINVAL, "upload: unknown buffer %d", buf);
  if (count < 0 || (uint64_t)count > b->mMax) return fx::fail(FX_ERANGE, "upload: %lld elements exceed %llu",
                                                             (long long)count, (unsigned long long)b->mMax);
  if (buf == FID) {
    // ids index the adjoint's target and gradients: only a whole permutation of 0..n-1 is taken
    if (count != mNumPoints)
      return fx::fail(FX_EINVAL, "upload: FID takes all %d ids at once, not %lld", mNumPoints, (long long)count);
    FX_TRY(CheckIds(static_cast<const uint32_t*>(host), count, "upload"));
  }
  FX_HIP(hipStreamSynchronize(m_Stream), "upload");
  FX_HIP(hipMemcpyHtoD(b->mGpu, const_cast<void*>(host), (size_t)count * b->mStride), "upload");
  return FX_OK;
}

int Particles::Download(int buf, void* host, int64_t count) {
  DataPtr* b = m_Points.GetBuffer(buf);
  if (b == nullptr || host == nullptr) return fx::fail(FX_EINVAL, "download: unknown buffer %d", buf);
  if (count < 0 || (uint64_t)count > b->mMax) return fx::fail(FX_ERANGE, "download: %lld elements exceed %llu",
                                                             (long long)count, (unsigned long long)b->mMax);
  FX_HIP(hipStreamSynchronize(m_Stream), "download");
  FX_HIP(hipMemcpyDtoH(host, b->mGpu, (size_t)count * b->mStride), "download");
  return FX_OK;
}

int Particles::DownloadAccel(int buf, void* host, int64_t count) {
  DataPtr* b = m_Accel.GetBuffer(buf);
  if (b == nullptr || host == nullptr) return fx::fail(FX_EINVAL, "download_accel: unknown buffer %d", buf);
  if (count < 0 || (uint64_t)count > b->mMax)
    return fx::fail(FX_ERANGE, "download_accel: %lld elements exceed %llu", (long long)count,
                    (unsigned long long)b->mMax);
  FX_HIP(hipStreamSynchronize(m_Stream), "download_accel");
  FX_HIP(hipMemcpyDtoH(host, b->mGpu, (size_t)count * b->mStride), "download_accel");
  return FX_OK;
}

int Particles::Scan(const uint32_t* in, uint32_t* out, int n) {
  // The role of the reference's TestPrefixSum (fluids4.0/fluid_system.cpp:2120-2160):
  // the grid scan applied to caller data, through temporary device buffers.
  if (m_Module == nullptr) return fx::fail(FX_ESTATE, "scan: call Initialize first");
  if (n <= 0 || in == nullptr || out == nullptr) return fx::fail(FX_EINVAL, "scan: bad arguments");
  const int n2 = n / SCAN_ELEMS + 1, n3 = n2 / SCAN_ELEMS + 1;
  if (n3 > SCAN_ELEMS) return fx::fail(FX_ERANGE, "scan: %d elements exceed 1024^3", n);
  void* b[6] = {};
  const size_t sz[6] = {(size_t)n, (size_t)n, (size_t)n2, (size_t)n2, (size_t)n3, (size_t)n3};
  int st = FX_OK;
  for (int k = 0; k < 6 && st == FX_OK; k++)
    if (hipMalloc(&b[k], sz[k] * sizeof(uint32_t)) != hipSuccess) st = fx::fail(FX_EHIP, "scan: hipMalloc");
  if (st == FX_OK && hipMemcpyHtoD(b[0], const_cast<uint32_t*>(in), (size_t)n * 4) != hipSuccess)
    st = fx::fail(FX_EHIP, "scan: upload");
  if (st == FX_OK) st = ScanLevels(b[0], b[1], b[2], b[3], b[4], b[5], n);
  if (st == FX_OK && hipStreamSynchronize(m_Stream) != hipSuccess) st = fx::fail(FX_EHIP, "scan: run");
  if (st == FX_OK && hipMemcpyDtoH(out, b[1], (size_t)n * 4) != hipSuccess) st = fx::fail(FX_EHIP, "scan: download");
  for (auto p : b)
    if (p) (void)hipFree(p);
  return st;
}

int Particles::ScanDeviceNoAlloc(const uint32_t* in, uint32_t* out, int n) {
  // Scan on the caller's device buffers with the upper levels in the
  // library's scratch array: no allocation, no host copy.
  if (m_Module == nullptr || m_ScanScratch == nullptr)
    return fx::fail(FX_ESTATE, "scan_device_noalloc: call Initialize first");
  if (n <= 0 || in == nullptr || out == nullptr || (const void*)in == (const void*)out)
    return fx::fail(FX_EINVAL, "scan_device_noalloc: bad arguments");
  if (n > FX_SCAN_NOALLOC_MAX)
    return fx::fail(FX_ERANGE, "scan_device_noalloc: %d elements exceed its scratch (at most %d)", n,
                    FX_SCAN_NOALLOC_MAX);
  uint32_t* s = static_cast<uint32_t*>(m_ScanScratch);
  FX_TRY(ScanLevels(const_cast<uint32_t*>(in), out, s, s + FX_SCAN_L2, s + 2 * FX_SCAN_L2,
                    s + 2 * FX_SCAN_L2 + FX_SCAN_L3, n));
  FX_HIP(hipStreamSynchronize(m_Stream), "scan_device_noalloc");
  return FX_OK;
}

int Particles::Mark(int tag) {
  if (m_Module == nullptr) return fx::fail(FX_ESTATE, "mark: call Initialize first");
  void* args[] = {&tag};
  return Launch(FUNC_MARK, 1, 64, args);
}

int Particles::Synchronize() {
  FX_HIP(hipStreamSynchronize(m_Stream), "synchronize");
  return FX_OK;
}

// ------------------------------------------------------------------- timing

int Particles::ProfileStages(int nsteps, double* ms_out) {
  if (nsteps <= 0 || ms_out == nullptr) return fx::fail(FX_EINVAL, "profile_stages: bad arguments");
  fx::Event evs[FX_NUM_STAGES + 1];
  hipEvent_t ev[FX_NUM_STAGES + 1];
  for (int k = 0; k <= FX_NUM_STAGES; k++) {
    FX_HIP(evs[k].create(), "profile_stages");
    ev[k] = evs[k].e;
  }
  for (int k = 0; k <= FX_NUM_STAGES; k++) ms_out[k] = 0.0;
  int st = FX_OK;
  for (int s = 0; s < nsteps && st == FX_OK; s++) {
    for (int k = 0; k < FX_NUM_STAGES && st == FX_OK; k++) {
      if (hipEventRecord(ev[k], m_Stream) != hipSuccess) st = fx::fail(FX_EHIP, "profile_stages: record");
      if (st == FX_OK) st = Stage(k);
    }
    if (st == FX_OK && hipEventRecord(ev[FX_NUM_STAGES], m_Stream) != hipSuccess)
      st = fx::fail(FX_EHIP, "profile_stages: record");
    if (st == FX_OK && hipEventSynchronize(ev[FX_NUM_STAGES]) != hipSuccess)
      st = fx::fail(FX_EHIP, "profile_stages: synchronize");
    for (int k = 0; k < FX_NUM_STAGES && st == FX_OK; k++) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, ev[k], ev[k + 1]) != hipSuccess) st = fx::fail(FX_EHIP, "profile_stages: elapsed");
      ms_out[k] += ms;
    }
    float tot = 0.0f;
    if (st == FX_OK && hipEventElapsedTime(&tot, ev[0], ev[FX_NUM_STAGES]) == hipSuccess)
      ms_out[FX_NUM_STAGES] += tot;
  }
  if (st != FX_OK) return st;
  for (int k = 0; k <= FX_NUM_STAGES; k++) ms_out[k] /= nsteps;
  return FX_OK;
}

int Particles::SampleParticles(float* brick, const int* res, const float* bmin, const float* bmax, float scalar) {
  if (mNumPoints == 0) return fx::fail(FX_ESTATE, "sample_particles: call Restart first");
  if (m_Batch) return fx::fail(FX_ESTATE, "sample_particles: one scene only");
  if (m_Done < FX_STAGE_SORT) return fx::fail(FX_ESTATE, "sample_particles: the cell table is made by a step's sort");
  if (brick == nullptr || res == nullptr || bmin == nullptr || bmax == nullptr)
    return fx::fail(FX_EINVAL, "sample_particles: null argument");
  if (res[0] <= 0 || res[1] <= 0 || res[2] <= 0 || (int64_t)res[0] * res[1] * res[2] > ((int64_t)1 << 31))
    return fx::fail(FX_ERANGE, "sample_particles: bad brick resolution %d x %d x %d", res[0], res[1], res[2]);
  const size_t nv = (size_t)res[0] * res[1] * res[2];
  void* d = nullptr;
  FX_HIP(hipMalloc(&d, nv * sizeof(float)), "sample_particles");
  fx_int3 r = {res[0], res[1], res[2]};
  fx_float3 lo = {bmin[0], bmin[1], bmin[2]}, hi = {bmax[0], bmax[1], bmax[2]};
  int n = mNumPoints;
  void* args[] = {&d, &r, &lo, &hi, &n, &scalar};
  const unsigned bx = 8, by = 8, bz = 4;
  hipError_t e = hipModuleLaunchKernel(m_Func[FUNC_SAMPLE], (unsigned)((res[0] + bx - 1) / bx),
                                       (unsigned)((res[1] + by - 1) / by), (unsigned)((res[2] + bz - 1) / bz), bx, by,
                                       bz, 0, m_Stream, args, nullptr);
  if (e == hipSuccess) e = hipStreamSynchronize(m_Stream);
  if (e == hipSuccess) e = hipMemcpyDtoH(brick, d, nv * sizeof(float));
  (void)hipFree(d);
  if (e != hipSuccess) return fx::hip_fail(e, "sample_particles", "sampleParticles");
  return FX_OK;
}

int Particles::EmitParticles(float frame, int emit) {
  // fluids4.0/fluid_system.cpp drives emitParticles once per frame with the
  // frame number; one scene only
  if (mNumPoints == 0) return fx::fail(FX_ESTATE, "emit_particles: call Restart first");
  if (m_Batch) return fx::fail(FX_ESTATE, "emit_particles: one scene only");
  if (emit < 0) return fx::fail(FX_EINVAL, "emit_particles: emit %d", emit);
  if (emit > mNumPoints) emit = mNumPoints;
  if (emit == 0) return FX_OK;
  ReleaseGraph();
  int n = mNumPoints;
  void* args[] = {&frame, &emit, &n};
  return Launch(FUNC_EMIT, (unsigned)((emit + 255) / 256), 256, args);
}

int Particles::MaxNbrWords(int* words) {
  if (mNumPoints == 0 || m_Stats == nullptr) return fx::fail(FX_ESTATE, "max_words: call Restart first");
  if (m_Done < FX_STAGE_PREFIX) return fx::fail(FX_ESTATE, "max_words: the counts and offsets are made by insert and scan");
  int four = 4;
  unsigned zero = 0;
  void* st = m_Stats;
  void* fargs[] = {&st, &zero, &four};
  FX_TRY(Launch(FUNC_FILL, 1, 64, fargs));
  if (m_Batch) {
    FX_TRY(BatchLaunch(FUNC_MAXWORDS_B, 0, FX_STAGE_PRESSURE, &st, nullptr, nullptr));
  } else {
    void* args[] = {&st, &mNumPoints};
    FX_TRY(LaunchP(FUNC_MAXWORDS, FX_STAGE_PRESSURE, args));
  }
  unsigned h = 0;
  FX_HIP(hipStreamSynchronize(m_Stream), "max_words");
  FX_HIP(hipMemcpyDtoH(&h, m_Stats, sizeof(h)), "max_words");
  *words = (int)h;
  return FX_OK;
}

int Particles::CountPairs(int64_t* cand, int64_t* pairs) {
  if (mNumPoints == 0 || m_Stats == nullptr) return fx::fail(FX_ESTATE, "count_pairs: call Restart first");
  if (m_Done < FX_STAGE_SORT) return fx::fail(FX_ESTATE, "count_pairs: the cell table is made by a step's sort");
  int four = 4;
  unsigned zero = 0;
  void* st = m_Stats;
  void* fargs[] = {&st, &zero, &four};
  FX_TRY(Launch(FUNC_FILL, 1, 64, fargs));
  if (m_Batch) {
    FX_TRY(BatchLaunch(FUNC_COUNT_PAIRS_B, 0, FX_STAGE_PRESSURE, &st, nullptr, nullptr));
  } else {
    void* args[] = {&st, &mNumPoints};
    FX_TRY(LaunchP(FUNC_COUNT_PAIRS, FX_STAGE_PRESSURE, args));
  }
  unsigned long long h[2] = {0, 0};
  FX_HIP(hipStreamSynchronize(m_Stream), "count_pairs");
  FX_HIP(hipMemcpyDtoH(h, m_Stats, sizeof(h)), "count_pairs");
  *cand = (int64_t)h[0];
  *pairs = (int64_t)h[1];
  return FX_OK;
}

int Particles::BenchSteps(int nsteps, bool use_graph, double* ms_total, double* stage_ms) {
  // nsteps steps with a HIP event before every stage and one after the last,
  // all on m_Stream; with use_graph the whole region (events included) is one
  // captured graph launched once.
  if (mNumPoints == 0) return fx::fail(FX_ESTATE, "bench: call Restart first");
  if (nsteps <= 0 || ms_total == nullptr || stage_ms == nullptr) return fx::fail(FX_EINVAL, "bench: bad arguments");
  FX_HIP(hipStreamSynchronize(m_Stream), "bench");
  const int nev = nsteps * FX_NUM_STAGES + 1;
  std::vector<hipEvent_t> ev(nev, nullptr);
  int st = FX_OK;
  for (int k = 0; k < nev && st == FX_OK; k++)
    if (hipEventCreate(&ev[k]) != hipSuccess) st = fx::fail(FX_EHIP, "bench: hipEventCreate");
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  const float time0 = m_Params.time;
  const int parity0 = m_Parity, done0 = m_Done;
  if (st == FX_OK && use_graph) {
    if (hipStreamBeginCapture(m_Stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      st = fx::fail(FX_EHIP, "bench: begin capture");
    else {
      m_Capturing = true;
      for (int s = 0; s < nsteps && st == FX_OK; s++)
        for (int k = 0; k < FX_NUM_STAGES && st == FX_OK; k++) {
          // an external event node, so that the replay records a timestamp
          const hipError_t er = hipEventRecordWithFlags(ev[s * FX_NUM_STAGES + k], m_Stream, hipEventRecordExternal);
          if (er != hipSuccess) st = fx::hip_fail(er, "bench", "event record in capture");
          if (st == FX_OK) st = (k == FX_STAGE_ADVANCE) ? Advance() : Stage(k);
        }
      if (st == FX_OK) {
        const hipError_t er = hipEventRecordWithFlags(ev[nev - 1], m_Stream, hipEventRecordExternal);
        if (er != hipSuccess) st = fx::hip_fail(er, "bench", "last event record in capture");
      }
      m_Capturing = false;
      hipError_t e = hipStreamEndCapture(m_Stream, &g);
      if (st == FX_OK && e != hipSuccess) st = fx::hip_fail(e, "bench", "hipStreamEndCapture");
      if (st == FX_OK && hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess)
        st = fx::fail(FX_EHIP, "bench: instantiate");
      if (st == FX_OK) st = SetDeviceClock();
      if (st == FX_OK && hipGraphLaunch(ge, m_Stream) != hipSuccess) st = fx::fail(FX_EHIP, "bench: launch");
      if (st == FX_OK)
        for (int s = 0; s < nsteps; s++) AdvanceTime();
    }
  } else if (st == FX_OK) {
    for (int s = 0; s < nsteps && st == FX_OK; s++)
      for (int k = 0; k < FX_NUM_STAGES && st == FX_OK; k++) {
        if (hipEventRecord(ev[s * FX_NUM_STAGES + k], m_Stream) != hipSuccess)
          st = fx::fail(FX_EHIP, "bench: event record");
        if (st == FX_OK) st = Stage(k);
      }
    if (st == FX_OK && hipEventRecord(ev[nev - 1], m_Stream) != hipSuccess) st = fx::fail(FX_EHIP, "bench: event record");
  }
  if (st == FX_OK && hipStreamSynchronize(m_Stream) != hipSuccess) st = fx::fail(FX_EHIP, "bench: synchronize");
  if (st == FX_OK) {
    for (int k = 0; k < FX_NUM_STAGES; k++) stage_ms[k] = 0.0;
    for (int s = 0; s < nsteps && st == FX_OK; s++)
      for (int k = 0; k < FX_NUM_STAGES && st == FX_OK; k++) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ev[s * FX_NUM_STAGES + k], ev[s * FX_NUM_STAGES + k + 1]) != hipSuccess)
          st = fx::fail(FX_EHIP, "bench: elapsed");
        stage_ms[k] += ms;
      }
    float tot = 0.0f;
    if (st == FX_OK && hipEventElapsedTime(&tot, ev[0], ev[nev - 1]) != hipSuccess) st = fx::fail(FX_EHIP, "bench: elapsed");
    *ms_total = tot;
    for (int k = 0; k < FX_NUM_STAGES; k++) stage_ms[k] /= nsteps;
  }
  if (st != FX_OK && use_graph) {
    // nothing of a failed capture ran: restore the host view of the sets
    m_Params.time = time0;
    if (m_Parity != parity0) {
      (void)m_Points.SwapGPU(&m_PointsTemp);
      m_Parity = parity0;
    }
    m_Done = done0;
  }
  if (ge) (void)hipGraphExecDestroy(ge);
  if (g) (void)hipGraphDestroy(g);
  for (auto e : ev)
    if (e) (void)hipEventDestroy(e);
  return st;
}

int Particles::TimeSteps(int nsteps, bool use_graph, double* ms_total) {
  if (nsteps < 0 || ms_total == nullptr) return fx::fail(FX_EINVAL, "time_steps: bad arguments");
  if (use_graph && (m_GraphExec == nullptr || m_GraphParity != m_Parity)) FX_TRY(CaptureGraph());
  if (use_graph) {
    // the device clock must hold the host time before the timed region
    FX_HIP(hipStreamSynchronize(m_Stream), "time_steps");
    FX_TRY(SetDeviceClock());
  }
  fx::Event ea, eb;
  FX_HIP(ea.create(), "time_steps");
  FX_HIP(eb.create(), "time_steps");
  hipEvent_t a = ea.e, b = eb.e;
  int st = FX_OK;
  if (hipEventRecord(a, m_Stream) != hipSuccess) st = fx::fail(FX_EHIP, "time_steps: record");
  if (st == FX_OK) {
    if (use_graph) {
      for (int k = 0; k < nsteps / 2 && st == FX_OK; k++) {
        if (hipGraphLaunch(m_GraphExec, m_Stream) != hipSuccess) st = fx::fail(FX_EHIP, "time_steps: graph launch");
        AdvanceTime();
        AdvanceTime();
        m_Done = FX_STAGE_ADVANCE;
      }
      if (st == FX_OK && (nsteps & 1)) st = Run();
    } else {
      for (int k = 0; k < nsteps && st == FX_OK; k++) st = Run();
    }
  }
  if (st == FX_OK && hipEventRecord(b, m_Stream) != hipSuccess) st = fx::fail(FX_EHIP, "time_steps: record");
  if (st == FX_OK && hipEventSynchronize(b) != hipSuccess) st = fx::fail(FX_EHIP, "time_steps: synchronize");
  float ms = 0.0f;
  if (st == FX_OK && hipEventElapsedTime(&ms, a, b) != hipSuccess) st = fx::fail(FX_EHIP, "time_steps: elapsed");
  *ms_total = ms;
  return st;
}
