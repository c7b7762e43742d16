// autotune.cpp -- launch-geometry autotuner of the SPH step (no reference
// counterpart: the reference launches every kernel with 512 threads per
// block, fluids5.0/src/particles.cpp:417-418).
//
// Every stage kernel maps one particle to one lane, so the tunable of each
// stage is its block size; particles per thread is fixed at one (the density
// and force stages already batch work per block through their LDS tiles and
// per-lane pair rings).  For each candidate block size the tuner profiles
// `nsteps` steps from the same saved state with every stage at that size,
// keeps for each stage the fastest size, checks that one step under the
// chosen sizes is bit-identical to one step under the previous sizes, and
// restores the state it started from.
#include <cstring>
#include <vector>

#include "fx_common.hpp"
#include "particles.hpp"

namespace {
const int kCandidates[] = {64, 128, 256, 512, 1024};
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
  int saved[FX_NUM_STAGES];
  memcpy(saved, m_StageThreads, sizeof(saved));
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
  auto set_all = [&](const int* t) -> int {
    memcpy(m_StageThreads, t, sizeof(m_StageThreads));
    return BuildBlockTables();
  };
  // one step under the given sizes from the saved state, downloaded
  auto one_step = [&](const int* t, std::vector<char>& out) -> int {
    FX_TRY(restore());
    FX_TRY(set_all(t));
    FX_TRY(Run());
    out.resize(36 * n);
    FX_HIP(hipStreamSynchronize(m_Stream), "autotune");
    FX_HIP(hipMemcpyDtoH(out.data(), m_Points.gpu(FPOS), 12 * n), "autotune");
    FX_HIP(hipMemcpyDtoH(out.data() + 12 * n, m_Points.gpu(FVEL), 12 * n), "autotune");
    FX_HIP(hipMemcpyDtoH(out.data() + 24 * n, m_Points.gpu(FVEVAL), 12 * n), "autotune");
    return FX_OK;
  };

  int st = save();
  int best[FX_NUM_STAGES];
  double best_ms[FX_NUM_STAGES];
  for (int k = 0; k < FX_NUM_STAGES; k++) {
    best[k] = saved[k];
    best_ms[k] = 1e300;
  }
  for (int T : kCandidates) {
    if (st != FX_OK) break;
    int all[FX_NUM_STAGES];
    for (int k = 0; k < FX_NUM_STAGES; k++) all[k] = T;
    double ms[FX_NUM_STAGES + 1];
    st = restore();
    if (st == FX_OK) st = set_all(all);
    if (st == FX_OK) st = ProfileStages(nsteps, ms);
    for (int k = 0; k < FX_NUM_STAGES && st == FX_OK; k++)
      if (ms[k] < best_ms[k]) {
        best_ms[k] = ms[k];
        best[k] = T;
      }
  }
  std::vector<char> a, b;
  if (st == FX_OK) st = one_step(saved, a);
  if (st == FX_OK) st = one_step(best, b);
  if (st == FX_OK && memcmp(a.data(), b.data(), a.size()) != 0)
    st = fx::fail(FX_EMISMATCH, "autotune: a step under the tuned block sizes differs from the previous sizes");
  // leave the system at its starting state with the tuned sizes (or the old
  // ones after a failure)
  const int rs = restore();
  const int ss = set_all(st == FX_OK ? best : saved);
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
