// capi.cpp -- the extern "C" entry points declared in include/fluidsx.h.
#include <new>
#include <string>

#include "fluidsx.h"
#include "fx_common.hpp"
#include "fx_scene.hpp"
#include "particles.hpp"

struct fx_system {
  Particles p;
};

namespace fx {
std::string& last_error() {
  static thread_local std::string msg;
  return msg;
}
}  // namespace fx

#define FX_CHECK_SYS(s) \
  if ((s) == nullptr) return fx::fail(FX_EINVAL, "%s: null fx_system", __func__)

extern "C" {

int fx_scene_derive(const fx_scene* scene, FParams_t* out) {
  if (scene == nullptr || out == nullptr) return fx::fail(FX_EINVAL, "fx_scene_derive: null argument");
  return fx::scene_derive(*scene, *out);
}

float fx_scene_spacing(const fx_scene* scene) {
  return scene == nullptr ? 0.0f : fx::scene_spacing(*scene);
}

int fx_scene_fill(const fx_scene* scene, float* pos, uint32_t* clr, int32_t n) {
  if (scene == nullptr || pos == nullptr) return fx::fail(FX_EINVAL, "fx_scene_fill: null argument");
  return fx::scene_fill(*scene, pos, clr, n);
}

int fx_create(const char* path, fx_system** out) {
  if (path == nullptr || out == nullptr) return fx::fail(FX_EINVAL, "fx_create: null argument");
  *out = nullptr;
  fx_system* s = new (std::nothrow) fx_system;
  if (s == nullptr) return fx::fail(FX_ERANGE, "fx_create: out of host memory");
  int st = s->p.Initialize(path);
  if (st != FX_OK) {
    delete s;
    return st;
  }
  *out = s;
  return FX_OK;
}

void fx_destroy(fx_system* s) { delete s; }

int fx_restart(fx_system* s, const fx_scene* scene) {
  FX_CHECK_SYS(s);
  if (scene == nullptr) return fx::fail(FX_EINVAL, "fx_restart: null scene");
  return s->p.Restart(*scene);
}

int fx_restart_batch(fx_system* s, const fx_scene* scenes, int32_t count) {
  FX_CHECK_SYS(s);
  return s->p.RestartBatch(scenes, count);
}

int32_t fx_num_scenes(const fx_system* s) { return s == nullptr ? 0 : s->p.NumScenes(); }

int fx_scene_range(const fx_system* s, int32_t scene, int32_t* first_particle, int32_t* num_particles,
                   int32_t* first_cell, int32_t* num_cells) {
  FX_CHECK_SYS(s);
  if (!first_particle || !num_particles || !first_cell || !num_cells)
    return fx::fail(FX_EINVAL, "fx_scene_range: null argument");
  return s->p.SceneRange(scene, first_particle, num_particles, first_cell, num_cells);
}

int fx_scene_params(const fx_system* s, int32_t scene, FParams_t* out) {
  FX_CHECK_SYS(s);
  if (out == nullptr) return fx::fail(FX_EINVAL, "fx_scene_params: null argument");
  return s->p.SceneParams(scene, out);
}

int fx_get_params(const fx_system* s, FParams_t* out) {
  FX_CHECK_SYS(s);
  if (out == nullptr) return fx::fail(FX_EINVAL, "fx_get_params: null argument");
  *out = s->p.Params();
  return FX_OK;
}

int fx_set_params(fx_system* s, const FParams_t* p) {
  FX_CHECK_SYS(s);
  if (p == nullptr) return fx::fail(FX_EINVAL, "fx_set_params: null argument");
  return s->p.SetParams(*p);
}

int32_t fx_num_points(const fx_system* s) { return s == nullptr ? 0 : s->p.NumPoints(); }
int32_t fx_grid_total(const fx_system* s) { return s == nullptr ? 0 : s->p.Params().gridTotal; }
int32_t fx_cell_table_size(const fx_system* s) { return s == nullptr ? 0 : s->p.CellTableSize(); }

int32_t fx_buffer_stride(int32_t buf) {
  switch (buf) {
    case FPOS: case FVEL: case FVEVAL: case FFORCE: return 12;
    case FCLR: case FPRESS: case FGCELL: case FGNDX: case FID: return 4;
    default: return 0;
  }
}

int fx_upload(fx_system* s, int32_t buf, const void* host, int64_t count) {
  FX_CHECK_SYS(s);
  return s->p.Upload(buf, host, count);
}

int fx_download(fx_system* s, int32_t buf, void* host, int64_t count) {
  FX_CHECK_SYS(s);
  return s->p.Download(buf, host, count);
}

int fx_download_accel(fx_system* s, int32_t buf, void* host, int64_t count) {
  FX_CHECK_SYS(s);
  return s->p.DownloadAccel(buf, host, count);
}

int fx_sample_particles(fx_system* s, float* brick, const int32_t* res, const float* bmin, const float* bmax,
                        float scalar) {
  if (s == nullptr) return fx::fail(FX_EINVAL, "null system");
  return s->p.SampleParticles(brick, res, bmin, bmax, scalar);
}

int fx_emit_particles(fx_system* s, float frame, int32_t emit) {
  if (s == nullptr) return fx::fail(FX_EINVAL, "null system");
  return s->p.EmitParticles(frame, emit);
}

int fx_save_points(fx_system* s, const char* path) {
  if (s == nullptr) return fx::fail(FX_EINVAL, "null system");
  return s->p.SavePoints(path);
}

int fx_load_points(fx_system* s, const char* path) {
  if (s == nullptr) return fx::fail(FX_EINVAL, "null system");
  return s->p.LoadPoints(path);
}

int fx_stage(fx_system* s, int32_t stage) {
  FX_CHECK_SYS(s);
  return s->p.Stage(stage);
}

int fx_run(fx_system* s, int32_t nsteps) {
  FX_CHECK_SYS(s);
  if (nsteps < 0) return fx::fail(FX_EINVAL, "fx_run: nsteps < 0");
  for (int k = 0; k < nsteps; k++) FX_TRY(s->p.Run());
  return FX_OK;
}

int fx_synchronize(fx_system* s) {
  FX_CHECK_SYS(s);
  return s->p.Synchronize();
}

int fx_scan(fx_system* s, const uint32_t* in, uint32_t* out, int32_t n) {
  FX_CHECK_SYS(s);
  return s->p.Scan(in, out, n);
}

int fx_set_stage_ppt(fx_system* s, int32_t stage, int32_t ppt) {
  FX_CHECK_SYS(s);
  return s->p.SetStagePpt(stage, ppt);
}

int32_t fx_stage_ppt(const fx_system* s, int32_t stage) {
  if (s == nullptr || stage < 0 || stage >= FX_NUM_STAGES) return 0;
  return s->p.StagePpt(stage);
}

int fx_set_persistent(fx_system* s, int32_t on) {
  FX_CHECK_SYS(s);
  return s->p.SetPersistent(on != 0);
}

int fx_mark(fx_system* s, int32_t tag) {
  FX_CHECK_SYS(s);
  return s->p.Mark(tag);
}

int fx_scan_device_noalloc(fx_system* s, const uint32_t* in, uint32_t* out, int32_t n) {
  FX_CHECK_SYS(s);
  return s->p.ScanDeviceNoAlloc(in, out, n);
}

int fx_graph_capture(fx_system* s) {
  FX_CHECK_SYS(s);
  return s->p.CaptureGraph();
}

int fx_graph_launch(fx_system* s, int32_t nlaunch) {
  FX_CHECK_SYS(s);
  if (nlaunch < 0) return fx::fail(FX_EINVAL, "fx_graph_launch: nlaunch < 0");
  return s->p.LaunchGraph(nlaunch);
}

int fx_profile_stages(fx_system* s, int32_t nsteps, double* ms_out) {
  FX_CHECK_SYS(s);
  return s->p.ProfileStages(nsteps, ms_out);
}

int fx_time_steps(fx_system* s, int32_t nsteps, int32_t use_graph, double* ms_total) {
  FX_CHECK_SYS(s);
  return s->p.TimeSteps(nsteps, use_graph != 0, ms_total);
}

int fx_bench_steps(fx_system* s, int32_t nsteps, int32_t use_graph, double* ms_total, double* stage_ms) {
  FX_CHECK_SYS(s);
  return s->p.BenchSteps(nsteps, use_graph != 0, ms_total, stage_ms);
}

int fx_count_pairs(fx_system* s, int64_t* candidates, int64_t* pairs) {
  FX_CHECK_SYS(s);
  if (candidates == nullptr || pairs == nullptr) return fx::fail(FX_EINVAL, "fx_count_pairs: null argument");
  return s->p.CountPairs(candidates, pairs);
}

int fx_set_stage_block_size(fx_system* s, int32_t stage, int32_t threads) {
  FX_CHECK_SYS(s);
  return s->p.SetStageBlockSize(stage, threads);
}

int32_t fx_stage_block_size(const fx_system* s, int32_t stage) {
  if (s == nullptr || stage < 0 || stage >= FX_NUM_STAGES) return 0;
  return s->p.StageBlockSize(stage);
}

int fx_set_block_size(fx_system* s, int32_t threads) {
  FX_CHECK_SYS(s);
  return s->p.SetBlockSize(threads);
}

int fx_autotune(fx_system* s, int32_t nsteps, int32_t* threads_out, double* ms_out) {
  if (s == nullptr) return fx::fail(FX_EINVAL, "null system");
  return s->p.Autotune(nsteps, threads_out, ms_out);
}

int fx_adjoint_configure(fx_system* s, const char* path, int32_t nsteps, int32_t every) {
  if (s == nullptr) return fx::fail(FX_EINVAL, "null system");
  return s->p.AdjConfigure(path, nsteps, every);
}

int fx_adjoint_set_target(fx_system* s, const float* target, int64_t count) {
  if (s == nullptr) return fx::fail(FX_EINVAL, "null system");
  return s->p.AdjSetTarget(target, count);
}

int fx_adjoint_forward(fx_system* s, double* loss, double* ms) {
  if (s == nullptr) return fx::fail(FX_EINVAL, "null system");
  return s->p.AdjForward(loss, ms);
}

int fx_adjoint_backward(fx_system* s, float* gpos, float* gvel, float* gveval, double* gvisc, double* gk,
                        double* ms) {
  if (s == nullptr) return fx::fail(FX_EINVAL, "null system");
  return s->p.AdjBackward(gpos, gvel, gveval, gvisc, gk, ms);
}

int64_t fx_adjoint_memory(const fx_system* s) { return s == nullptr ? 0 : s->p.AdjMemory(); }

int fx_adjoint_segment(const fx_system* s, int32_t* every) {
  if (s == nullptr || every == nullptr) return fx::fail(FX_EINVAL, "null argument");
  *every = s->p.AdjSegment();
  return FX_OK;
}

const char* fx_last_error(void) { return fx::last_error().c_str(); }

const char* fx_version(void) { return "fluidsx 0.1 gfx950"; }

}  // extern "C"
