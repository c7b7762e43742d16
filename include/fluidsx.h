/*
 * fluidsx.h -- C ABI of the MI355X SPH step (libfluidsx.so).
 *
 * Plain pointers and sizes only.  Every call returns 0 on success or a
 * negative FX_E* code; fx_last_error() then holds a message.  The reference
 * reports errors by printing and calling exit() (cuCheck,
 * fluids5.0/helpers/common_defs.cpp:223-252; DataX exit(-9/-11/-12),
 * fluids5.0/src/datax.cpp:76-79,114-117, dataptr.cpp:280-283); here they are
 * status codes.
 *
 * Threading: like the reference (one host thread, one context), an
 * fx_system is not thread-safe.  All device work of one fx_system is ordered
 * on one HIP stream owned by it (default priority); the reference uses the
 * NULL stream.
 */
#ifndef FLUIDSX_H
#define FLUIDSX_H

#include <stdint.h>

#include "fluidsx_params.h"

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define FX_OK 0
#define FX_EINVAL (-1)   /* bad argument */
#define FX_EHIP (-2)     /* a HIP runtime call failed */
#define FX_ESTATE (-3)   /* call not valid in the current state */
#define FX_ERANGE (-4)   /* size beyond a documented limit */
#define FX_EMISMATCH (-5) /* buffer layouts do not match (DataX exit -9/-11/-12) */

/* per-function stages of one step, in Particles::Run order
 * (fluids5.0/src/particles.cpp:382-394) */
#define FX_STAGE_INSERT 0   /* Particles::InsertParticles, particles.cpp:716-730 */
#define FX_STAGE_PREFIX 1   /* Particles::PrefixScanParticles, particles.cpp:774-815 */
#define FX_STAGE_SORT 2     /* Particles::CountingSort, particles.cpp:832-842 */
#define FX_STAGE_PRESSURE 3 /* Particles::ComputePressure, particles.cpp:852-857 */
#define FX_STAGE_FORCE 4    /* Particles::ComputeForce, particles.cpp:914-919 */
#define FX_STAGE_ADVANCE 5  /* Particles::Advance, particles.cpp:413-418 */
#define FX_NUM_STAGES 6

/*
 * Scene description: what SetupDefaultParams / SetupExampleParams /
 * AddPointsInVolume hard-code in the reference (particles.cpp:122-263,
 * 318-379), as data.
 */
typedef struct fx_scene {
  int32_t num_points;     /* N = mMaxPoints = mNumPoints (particles.cpp:95-119) */
  int32_t reserved0;
  uint64_t seed;          /* counter-based jitter RNG key */
  float dt, sim_scale, smooth_radius, mass, rest_density, visc, int_stiff;
  float bound_stiff, bound_damp, accel_limit, vel_limit;
  float bound_slope, bound_friction, bound_wall_force, bound_wall_freq;
  float pradius;      /* boundary contact radius (m) */
  float grid_density; /* cell edge = 1.5 h / grid_density (particles.cpp:551) */
  float grav_amt;
  fx_float3 grav_dir;
  fx_float3 bound_min, bound_max, init_min, init_max;
  float spacing; /* lattice spacing (world); 0 = rest spacing, shrunk to fit N */
  float jitter;  /* uniform jitter amplitude as a fraction of the spacing */
} fx_scene;

typedef struct fx_system fx_system;

/* ---- host-only set-up (no GPU needed) ---- */

/* Derives the full FParams_t of a scene: defaults, grid
 * (Particles::RebuildAccelGrid, particles.cpp:548-626), kernel constants
 * (Particles::UpdateParams, particles.cpp:657-679), launch geometry. */
int fx_scene_derive(const fx_scene* scene, FParams_t* out);

/* Lattice spacing the scene uses (rest spacing, or the largest spacing whose
 * lattice in [init_min, init_max] holds num_points sites). */
float fx_scene_spacing(const fx_scene* scene);

/* Initial particle lattice (Particles::AddPointsInVolume,
 * particles.cpp:318-379): y-layers bottom-up, x fastest then z, seeded
 * uniform jitter, stop at n.  pos holds 3n floats, clr n uints.  Returns the
 * number of particles written. */
int fx_scene_fill(const fx_scene* scene, float* pos, uint32_t* clr, int32_t n);

/* Scene files in the reference's scene.xml schema (its README.md "Scene
 * parameters", fluids5.0/assets/scene.xml): a <Fluid> section, then
 * <Scene id="k"> sections; a scene's keys override the Fluid section's, as
 * FluidSystem::ParseXML applies "Fluid" and then the scene
 * (fluids3.2/fluids/fluid_system.cpp:2268-2309).  text is the file's
 * contents, NUL-terminated.
 * fx_scene_from_xml: *scene holds the defaults on entry (fields the file does
 * not set keep them) and the scene on return.  name (name_len bytes) receives
 * <Name>; unused (unused_len bytes) receives "Key=value\n" for each key of a
 * force or viewer setting the step does not have (PointGravAmt, PointGravPos,
 * WaveForceMax, SearchDist, Draw*), or FX_ERANGE if it does not fit; either
 * may be NULL.  FX_EINVAL for malformed text, an unknown key, a value that
 * is not a number or three numbers, a missing scene or a <Num> that is not
 * positive (*scene unchanged then).
 * fx_scene_xml_ids: the ids of the <Scene> sections in file order (count of
 * them in *count; FX_ERANGE if more than max_ids). */
int fx_scene_from_xml(const char* text, int32_t scene_id, fx_scene* scene, char* name, int32_t name_len,
                      char* unused, int32_t unused_len);
int fx_scene_xml_ids(const char* text, int32_t* ids, int32_t max_ids, int32_t* count);

/* ---- lifecycle ---- */

/* Particles::Initialize (particles.cpp:68-93): loads the code object by path,
 * resolves the kernels and the FParams/FPnts/FPntTmp/FAccel symbols. */
int fx_create(const char* code_object_path, fx_system** out);
/* ~Particles (particles.cpp:44-51). */
void fx_destroy(fx_system* sys);

/* Particles::Restart (particles.cpp:95-119): derive params, allocate the
 * particle and grid buffers, fill the lattice, upload everything. */
int fx_restart(fx_system* sys, const fx_scene* scene);

/* A batch of independent scenes in one system (no reference counterpart:
 * the reference simulates one scene per Particles object).  The scenes'
 * particles are concatenated in the particle buffers (scene s occupies
 * [first_particle, first_particle + num_particles)), their cell tables
 * (each with its out-of-range slot) are concatenated in AGRIDCNT/AGRIDOFF,
 * and every stage is one launch over all scenes.  Each scene keeps its own
 * box, grid and parameters; all share the time step count. */
int fx_restart_batch(fx_system* sys, const fx_scene* scenes, int32_t count);
int32_t fx_num_scenes(const fx_system* sys);
int fx_scene_range(const fx_system* sys, int32_t scene, int32_t* first_particle, int32_t* num_particles,
                   int32_t* first_cell, int32_t* num_cells);
int fx_scene_params(const fx_system* sys, int32_t scene, FParams_t* out);

/* FParams_t currently in use (host copy). */
int fx_get_params(const fx_system* sys, FParams_t* out);
/* Particles::UpdateParams (particles.cpp:657-679): takes the input fields of
 * p (rates, stiffness, walls, time ...), re-derives the kernel constants and
 * uploads the 688-byte block.  The grid must not change (FX_ESTATE). */
int fx_set_params(fx_system* sys, const FParams_t* p);

int32_t fx_num_points(const fx_system* sys);
int32_t fx_grid_total(const fx_system* sys);
/* Entries of AGRIDCNT / AGRIDOFF: gridTotal + 1, or the sum over a batch. */
int32_t fx_cell_table_size(const fx_system* sys);

/* ---- buffers (DataX::Commit / Retrieve, datax.cpp:124-137) ---- */

/* Element size of particle buffer id (FPOS..FID), 0 for an unknown id. */
int32_t fx_buffer_stride(int32_t buf);
/* Copy count elements of particle buffer buf between host and device.
 * Synchronous with respect to the system's stream.  An upload of FID must
 * hold all n ids and be a permutation of 0..n-1 (FX_EINVAL otherwise): the
 * ids index the adjoint's target and gradients. */
int fx_upload(fx_system* sys, int32_t buf, const void* host, int64_t count);
int fx_download(fx_system* sys, int32_t buf, void* host, int64_t count);
/* Grid buffers AGRID (N uints), AGRIDCNT / AGRIDOFF (gridTotal+1 uints). */
int fx_download_accel(fx_system* sys, int32_t buf, void* host, int64_t count);

/* Particle state files in the Fluids 4.0 .pts layout
 * (FluidSystem::SavePoints / RunPlayback, fluids4.0/fluid_system.cpp:1511-1545,
 * 1787-1843): int32 count, int32 fields, then per field int32 type
 * (0 char, 1 int, 2 float), int32 channels and the values.  Written: pos,
 * vel, clr (the 4.0 fields), then veval and id.  Loading needs the same
 * particle count as the system and accepts 4.0 files (veval = vel,
 * ids 0..n-1 then). */
int fx_save_points(fx_system* sys, const char* path);
int fx_load_points(fx_system* sys, const char* path);

/* ---- stepping ---- */

/* One per-function stage (FX_STAGE_*), enqueued on the stream.  A stage whose
 * input tables the current step has not made yet is refused (FX_ESTATE):
 * prefix needs the insert, sort needs the prefix scan right after it,
 * pressure needs the sort, force needs the pressure stage; insert and advance
 * can always run. */
int fx_stage(fx_system* sys, int32_t stage);
/* Particles::Run n times (stages in order, then AdvanceTime). */
int fx_run(fx_system* sys, int32_t nsteps);
int fx_synchronize(fx_system* sys);
/* Enqueues one dispatch of the no-op kernel fxMark(tag): a point of the
 * stream that kernel traces and counter runs show (no reference
 * counterpart; the reference marks host time with PERF_PUSH/POP,
 * fluids5.0/src/particles.cpp:384-389). */
int fx_mark(fx_system* sys, int32_t tag);

/* Exclusive prefix sum of n uints with the grid-offset scan kernels
 * (prefixSum / prefixFixup, up to three levels), through temporary device
 * buffers: the role of TestPrefixSum (fluids4.0/fluid_system.cpp:2120-2160).
 * n <= 1024^3 - 1. */
int fx_scan(fx_system* sys, const uint32_t* in, uint32_t* out, int32_t n);

/* The same exclusive scan on device buffers, allocating nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  in and out are
 * distinct device buffers of n uint32 on the system's device; the caller's
 * writes to in must be complete.  The upper scan levels use a fixed device
 * array of the library (about 1 MB per GPU, held until the process ends), so
 * n <= FX_SCAN_NOALLOC_MAX (else FX_ERANGE), and calls that overlap on one GPU
 * race.  Returns when out is written. */
#define FX_SCAN_NOALLOC_MAX (1 << 27)
int fx_scan_device_noalloc(fx_system* sys, const uint32_t* in, uint32_t* out, int32_t n);

/* Captures two consecutive steps (the device sets alternate roles each step)
 * into a HIP graph; fx_graph_launch replays it, advancing 2*nlaunch steps.
 * The simulation time then lives on the device. */
int fx_graph_capture(fx_system* sys);
int fx_graph_launch(fx_system* sys, int32_t nlaunch);

/* Runs nsteps eagerly with HIP events around every stage on the system's
 * stream; ms_out[FX_NUM_STAGES] receives the mean milliseconds per stage and
 * ms_out[FX_NUM_STAGES] the mean step time. */
int fx_profile_stages(fx_system* sys, int32_t nsteps, double* ms_out);
/* Wall time of nsteps: graph replays when use_graph, else eager fx_run;
 * measured with HIP events on the system's stream.  */
int fx_time_steps(fx_system* sys, int32_t nsteps, int32_t use_graph, double* ms_total);

/* Benchmark region: nsteps steps with a HIP event before every stage and
 * after the last, all on the system's stream; with use_graph the whole region
 * (events included) is captured as one graph and launched once.  ms_total
 * receives the region's time, stage_ms[FX_NUM_STAGES] the mean per stage. */
int fx_bench_steps(fx_system* sys, int32_t nsteps, int32_t use_graph, double* ms_total, double* stage_ms);

/* Neighbour statistics of the current cell-sorted state (call between the
 * sort and the next advance): the sum over in-range particles of stencil
 * candidates and of in-radius pairs (0 < d^2 < rd2); the role of the
 * reference's neighbour-count validation (fluids4.0/fluid_system.cpp:959-1011). */
int fx_count_pairs(fx_system* sys, int64_t* candidates, int64_t* pairs);

/* Gaussian splat of the particles onto a res[0] x res[1] x res[2] brick over
 * [bmin, bmax] using the cell table (sampleParticles,
 * fluids4.0/fluid_kernels.cu:259-300): brick[(z*res[1] + y)*res[0] + x] =
 * scalar * sum of gausskern exp(-r^2 / (2 r2 / 8)) over the stencil's
 * particles with 0 < d^2 < rd2.  brick: host array of the voxel count.
 * One scene only. */
int fx_sample_particles(fx_system* sys, float* brick, const int32_t* res, const float* bmin, const float* bmax,
                        float scalar);
/* Respawns min(emit, particles) particles at points of the box
 * (-170,10,-20)..(-190,60,20) quantized to 1/256 per axis, velocity (5,-2,0),
 * zero veval and force (emitParticles, fluids4.0/fluid_kernels.cu:231-257).
 * The draws are a counter-based SplitMix64 of (frame, thread) and the
 * respawned particles a frame-seeded bijection of the thread index (distinct
 * particles, a result independent of thread order), where the reference uses
 * per-particle XORWOW states and index v & (n-1).  One scene only. */
int fx_emit_particles(fx_system* sys, float frame, int32_t emit);

/* Launch geometry: threads per block of the per-particle kernels (64..1024,
 * multiple of 64), for every stage or for one.  Results do not depend on it. */
int fx_set_block_size(fx_system* sys, int32_t threads);
int fx_set_stage_block_size(fx_system* sys, int32_t stage, int32_t threads);
int32_t fx_stage_block_size(const fx_system* sys, int32_t stage);
/* Particles per thread (1..8) of the streaming stages FX_STAGE_INSERT,
 * FX_STAGE_SORT and FX_STAGE_ADVANCE: a block then covers threads * ppt
 * consecutive particles (other stages: 1 only, FX_EINVAL otherwise).  The
 * code object reads it from its module constant FPpt, which a host that
 * does not set it leaves at 1.  Results do not depend on it. */
int fx_set_stage_ppt(fx_system* sys, int32_t stage, int32_t ppt);
int32_t fx_stage_ppt(const fx_system* sys, int32_t stage);
/* Batch stage launches (fx_restart_batch): persistent (1) runs each stage as
 * one launch of as many blocks as the device holds at once, pulling (scene,
 * block) work items from per-XCD queues (each XCD drains a contiguous range
 * of items, then helps the others); 0 (the default, faster on MI355X: the
 * claims cost more than the hardware dispatcher) launches one block per item.
 * Results do not depend on it. */
int fx_set_persistent(fx_system* sys, int32_t on);

/* Autotuner (no reference counterpart; the reference launches everything with
 * 512 threads per block, particles.cpp:417-418): profiles nsteps steps from
 * the current state for each block size in {64, 128, 256, 512, 1024} at one
 * particle per thread and keeps the fastest per stage, then, at those sizes,
 * 2 and 4 particles per thread for the streaming stages (fx_set_stage_ppt),
 * keeping what beats one; checks that a step under the chosen configuration
 * is bit-identical to a step under the previous one (FX_EMISMATCH
 * otherwise), and leaves the system at the state it started from with the
 * chosen configuration.  threads_out / ms_out (optional): FX_NUM_STAGES
 * entries each; the particles per thread chosen: fx_stage_ppt. */
int fx_autotune(fx_system* sys, int32_t nsteps, int32_t* threads_out, double* ms_out);

/* ---- reverse mode (no reference counterpart: the reference has no
 * differentiable path) ----
 *
 * Loss of an nsteps rollout from the current state, L_s = (1/n_s) sum over
 * scene s of |x_T - x*|^2 (FP64 reduction), and its gradient with respect to
 * the initial positions, velocities and evaluation velocities of every
 * particle (by particle id, FID) and to each scene's viscosity and internal
 * stiffness.  Adjoint kernels of density, force and advance run over the
 * recorded forward steps; neighbour sets, cells, wall and limiter branches
 * and the sort permutation of the forward are held fixed.  State is
 * checkpointed every `checkpoint_every` steps; the forward records the last
 * segment's steps and the backward recomputes each other segment once. */

/* Loads the adjoint code object (NULL: adjoint_gfx950.hsaco beside the step's
 * code object) and allocates checkpoints and tape for the current Restart.
 * It runs the insert and prefix stages on the current state (reordering
 * nothing) to size each tape slot's neighbour bitmask: twice the words the
 * densest stencil needs, 32 to FX_NBR_WORDS per particle; a particle needing
 * more during the rollout records no bits and is walked by testing its
 * candidates, with the same results.  checkpoint_every <= 0 chooses the
 * fewest segments, of equal length, whose tape and checkpoints fit in 60 % of
 * the device memory free at the call (the backward recomputes all segments
 * but the last; a rollout in one segment needs no recomputation).  When a
 * second tape buffer also fits, the backward recomputes each segment while
 * the adjoint of the next one runs. */
int fx_adjoint_configure(fx_system* sys, const char* code_object_path, int32_t nsteps,
                         int32_t checkpoint_every);
/* The segment length in use (the chosen one when configured with <= 0). */
int fx_adjoint_segment(const fx_system* sys, int32_t* checkpoint_every);
/* Target positions x*, 3 floats per particle in particle-id order (count =
 * particles).  Defaults to zero. */
int fx_adjoint_set_target(fx_system* sys, const float* target, int64_t count);
/* Runs the rollout (the system ends at step nsteps), records checkpoints and
 * the last segment's steps, writes the loss of every scene (fx_num_scenes
 * entries).  ms (optional) receives the GPU time of the call. */
int fx_adjoint_forward(fx_system* sys, double* loss, double* ms);
/* Gradients of the last fx_adjoint_forward's loss (each output optional):
 * grad_pos / grad_vel / grad_veval 3 floats per particle by id, grad_visc /
 * grad_stiff one double per scene.  The system is left at step nsteps. */
int fx_adjoint_backward(fx_system* sys, float* grad_pos, float* grad_vel, float* grad_veval, double* grad_visc,
                        double* grad_stiff, double* ms);
/* Device bytes held by checkpoints, tape and adjoint buffers. */
int64_t fx_adjoint_memory(const fx_system* sys);

/* Last error message of the calling thread. */
const char* fx_last_error(void);
const char* fx_version(void);

#ifdef __cplusplus
}
#endif

#endif /* FLUIDSX_H */
