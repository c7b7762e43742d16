// autotune.cpp -- launch-geometry autotuner of the SPH step (no reference
// counterpart: the reference launches every kernel with 512 threads per
// block, fluids5.0/src/particles.cpp:417-418).
//
// Two tunables: each stage's block size, and the particles per thread of the
// streaming stages (insert, sort, advance; the density and force stages
// already batch work per block through their LDS tiles and per wave through
// their word queues, one particle per lane).  For each candidate block size
// the tuner profiles `nsteps` steps from the same saved state with every
// stage at that size and one particle per thread, keeps for each stage the
// fastest size; then at those sizes it profiles 2 and 4 particles per thread
// for the streaming stages and keeps, per stage, what beats one.  It checks
// that one step under the chosen configuration is bit-identical to one step
// under the previous one, and restores the state it started from.
#include <cstring>
#include <vector>

#include "fx_common.hpp"
#include "particles.hpp"

namespace {
const int kCandidates[] = {64, 128, 256, 512, 1024};
const int kPptCandidates[] = {2, 4};
const bool kPptStage[FX_NUM_STAGES] = {true, false, true, false, false, true};
const int kSaved[5] = {FPOS, FVEL, FVEVAL, FCLR, FID};
const size_t kSavedBytes[5] = {12, 12, 12, 4, 4};
}  // namespace

int Particles::Autotune(int nsteps, int* chosen, double* stage_ms) {
  if (mNumPoints <= 0) return fx::fail(FX_ESTATE, "autotune: call Restart first");
  if (nsteps <= 0) return fx::fail(FX_EINVAL, "autotune: nsteps %d", nsteps);
  FX_HIP(hipStreamSynchronize(m_Stream), "autotune");
  ReleaseGraph();
  const size_t n = (size_t)mNumPoints;
  void* snap = nullptr;
  FX_HIP(hipMalloc(&snap, 44 * n), "autotune: snapshot");
  const float time0 = m_Params.time;
  int saved[FX_NUM_STAGES], saved_ppt[FX_NUM_STAGES];
  memcpy(saved, m_StageThreads, sizeof(saved));
  memcpy(saved_ppt, m_StagePpt, sizeof(saved_ppt));
  auto save = [&]() -> int {
    char* c = static_cast<char*>(snap);
    for (int b = 0; b < 5; b++) {
      FX_HIP(hipMemcpyAsync(c, m_Points.gpu(kSaved[b]), kSavedBytes[b] * n, hipMemcpyDeviceToDevice, m_Stream),
             "autotune");
      c += kSavedBytes[b] * n;
    }
    return FX_OK;
  };
  auto restore = [&]() -> int {
    const char* c = static_cast<const char*>(snap);
    for (int b = 0; b < 5; b++) {
      FX_HIP(hipMemcpyAsync(m_Points.gpu(kSaved[b]), c, kSavedBytes[b] * n, hipMemcpyDeviceToDevice, m_Stream),
             "autotune");
      c += kSavedBytes[b] * n;
    }
    m_Params.time = time0;
    return FX_OK;
  };
  auto set_all = [&](const int* t, const int* ppt) -> int {
    memcpy(m_StageThreads, t, sizeof(m_StageThreads));
    memcpy(m_StagePpt, ppt, sizeof(m_StagePpt));
    FX_TRY(UploadPpt());
    return BuildBlockTables();
  };
  // one step under the given configuration from the saved state, downloaded
  auto one_step = [&](const int* t, const int* ppt, std::vector<char>& out) -> int {
    FX_TRY(restore());
    FX_TRY(set_all(t, ppt));
    FX_TRY(Run());
    out.resize(36 * n);
    FX_HIP(hipStreamSynchronize(m_Stream), "autotune");
    FX_HIP(hipMemcpyDtoH(out.data(), m_Points.gpu(FPOS), 12 * n), "autotune");
    FX_HIP(hipMemcpyDtoH(out.data() + 12 * n, m_Points.gpu(FVEL), 12 * n), "autotune");
    FX_HIP(hipMemcpyDtoH(out.data() + 24 * n, m_Points.gpu(FVEVAL), 12 * n), "autotune");
    return FX_OK;
  };

  int st = save();
  int best[FX_NUM_STAGES], best_ppt[FX_NUM_STAGES], one[FX_NUM_STAGES];
  double best_ms[FX_NUM_STAGES];
  for (int k = 0; k < FX_NUM_STAGES; k++) {
    best[k] = saved[k];
    best_ppt[k] = 1;
    one[k] = 1;
    best_ms[k] = 1e300;
  }
  for (int T : kCandidates) {
    if (st != FX_OK) break;
    int all[FX_NUM_STAGES];
    for (int k = 0; k < FX_NUM_STAGES; k++) all[k] = T;
    double ms[FX_NUM_STAGES + 1];
    st = restore();
    if (st == FX_OK) st = set_all(all, one);
    if (st == FX_OK) st = ProfileStages(nsteps, ms);
    for (int k = 0; k < FX_NUM_STAGES && st == FX_OK; k++)
      if (ms[k] < best_ms[k]) {
        best_ms[k] = ms[k];
        best[k] = T;
      }
  }
  for (int ppt : kPptCandidates) {
    if (st != FX_OK) break;
    int cand[FX_NUM_STAGES];
    for (int k = 0; k < FX_NUM_STAGES; k++) cand[k] = kPptStage[k] ? ppt : 1;
    double ms[FX_NUM_STAGES + 1];
    st = restore();
    if (st == FX_OK) st = set_all(best, cand);
    if (st == FX_OK) st = ProfileStages(nsteps, ms);
    for (int k = 0; k < FX_NUM_STAGES && st == FX_OK; k++)
      if (kPptStage[k] && ms[k] < best_ms[k]) {
        best_ms[k] = ms[k];
        best_ppt[k] = ppt;
      }
  }
  std::vector<char> a, b;
  if (st == FX_OK) st = one_step(saved, saved_ppt, a);
  if (st == FX_OK) st = one_step(best, best_ppt, b);
  if (st == FX_OK && memcmp(a.data(), b.data(), a.size()) != 0)
    st = fx::fail(FX_EMISMATCH, "autotune: a step under the tuned configuration differs from the previous one");
  // leave the system at its starting state with the tuned configuration (or
  // the old one after a failure)
  const int rs = restore();
  const int ss = st == FX_OK ? set_all(best, best_ppt) : set_all(saved, saved_ppt);
  const hipError_t se = hipStreamSynchronize(m_Stream);
  (void)hipFree(snap);
  if (st != FX_OK) return st;
  if (rs != FX_OK) return rs;
  if (ss != FX_OK) return ss;
  if (se != hipSuccess) return fx::hip_fail(se, "autotune", "hipStreamSynchronize");
  for (int k = 0; k < FX_NUM_STAGES; k++) {
    if (chosen) chosen[k] = best[k];
    if (stage_ms) stage_ms[k] = best_ms[k];
  }
  return FX_OK;
}
