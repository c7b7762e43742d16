/*
 * fluidsx_params.h -- plain-C layout of the SPH step's shared state.
 *
 * Shared by the gfx950 kernels, the host library and the CPU oracle.
 * Every type here is a POD whose byte layout equals the reference's:
 *
 *   - buffer ids FPOS..FGNDX and AGRID..AAUXSCAN2, GRID_UNDEF
 *       (fluids5.0/src/fluid.h:12-31)
 *   - FParams_t, 688 bytes, with 12-byte float3/int3 members
 *       (fluids5.0/src/fluid.h:45-82; host Vector3DF is layout-identical,
 *        fluids5.0/helpers/vec.h:284)
 *   - cuDataX, the 16-pointer device table of one DataX
 *       (fluids5.0/src/datax.h:37-47)
 *
 * Additions (no counterpart in the reference, all in slots the reference
 * leaves unused):
 *   - FID = 8: a uint32 particle id carried through the counting sort so that
 *     results can be compared per particle.
 *   - ANBRBITS, AFREC: neighbour bitmask and packed per-particle record that
 *     the density stage leaves for the force stage.
 *   - ASLABSTART: where each x slab of a cell begins, written by the full
 *     sort for the density stage, which skips the slabs of a stencil row that
 *     lie beyond a particle's radius.
 *   - ATAILFLAG .. ATAILAUXS2: out-of-range flags and their exclusive scan,
 *     which give the tail bucket a stable (particle-order) ranking however
 *     many particles it holds.
 *   - AGRIDCNT / AGRIDOFF hold gridTotal+1 entries.  Slot gridTotal counts the
 *     particles whose cell is GRID_UNDEF, so the scan places them in a
 *     deterministic tail after every valid cell instead of leaving stale data
 *     there (fluids5.0/src/particles.cu:78 skips them).
 */
#ifndef FLUIDSX_PARAMS_H
#define FLUIDSX_PARAMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FX_GRID_UNDEF 2147483647 /* fluid.h:12 */
#define FX_EPSILON 0.00001f      /* particles.cuh:17, wall activation */

/* particle buffers (fluid.h:15-22) */
#define FPOS 0
#define FVEL 1
#define FCLR 2
#define FVEVAL 3
#define FFORCE 4
#define FPRESS 5 /* holds DENSITY, particles.cu:154-156 */
#define FGCELL 6
#define FGNDX 7
#define FID 8 /* addition: particle id */
#define FX_NUM_PBUF 9

/* acceleration-grid buffers (fluid.h:25-31) */
#define AGRID 0
#define AGRIDCNT 1
#define AGRIDOFF 2
#define AAUXARRAY1 3
#define AAUXSCAN1 4
#define AAUXARRAY2 5
#define AAUXSCAN2 6
/* additions, written by the density stage for the force stage (optional: a
 * host that does not allocate them gets the force stage without them) */
#define ANBRBITS 7 /* neighbour bitmask: word w of particle i at [w * N + i] */
#define AFREC 8    /* force record: 2 float4 per particle, (pos, rho), (veval, 0) */
#define ATAILFLAG 9   /* 1 for a particle whose cell is GRID_UNDEF, else 0 (N entries) */
#define ATAILSCAN 10  /* exclusive scan of ATAILFLAG: a tail particle's stable rank */
#define ATAILAUX1 11  /* scan levels of ATAILFLAG */
#define ATAILAUXS1 12
#define ATAILAUX2 13
#define ATAILAUXS2 14
#define ASLABSTART 15 /* per cell c and x slab k: members of c in slabs below k, uint32 at [c * FX_SLABS + k] */
#define FX_NUM_ABUF 16
#define FX_NBR_WORDS 256 /* bitmask words per particle (8192 candidates) */
#define FX_SLABS 16      /* x slabs per cell: the in-cell order of the ranked sort */
#define FX_MEMBER_BITS 28 /* AGRID member entries: index in the low 28 bits, x slab above */

/* Module constant FHostCaps (int, default 0): what the host allocated and
 * runs beyond the reference's layout.  A host that binds the code object the
 * way Fluids 5.0 does (particles.cpp:267-300,596-626,716-919) leaves it 0:
 * AGRIDCNT/AGRIDOFF hold gridTotal entries, out-of-range particles are not
 * counted, and countingSortFull ranks every particle by its insert slot
 * (gndx) and moves the reference's eight buffers.  libfluidsx sets both bits. */
#define FX_CAP_TAIL 1   /* AGRIDCNT/AGRIDOFF hold gridTotal+1 entries (tail bucket) */
#define FX_CAP_RANKED 2 /* countingSortIndex runs before countingSortFull (fixed in-cell order: x slab,
                         * then particle index); countingSortFullRanked is that sort alone, the faster
                         * kernel.  AGRID then holds member entries (index | slab << FX_MEMBER_BITS), so
                         * the host keeps fewer than 2^FX_MEMBER_BITS particles */

#define FX_DMAXBUF 16 /* datax.h:37 */

typedef struct fx_float3 {
  float x, y, z;
} fx_float3;
typedef struct fx_int3 {
  int x, y, z;
} fx_int3;

/* FParams_t, member for member (fluid.h:45-82). */
typedef struct FParams_t {
  int mode, example;
  float time, dt, sim_scale;
  unsigned char capture; /* C++ bool in the reference: 1 byte */

  int numThreads, numBlocks;
  int gridThreads, gridBlocks;
  int szPnts, szHash, szGrid;
  int stride, pnum, chk;

  float pdist, pmass, prest_dens, pintstiff;
  float pradius, psmoothradius, pspacing, r2, pvisc;

  float AL, AL2, VL, VL2;
  float d2, rd2, vterm;
  float poly6kern, spikykern, lapkern, gausskern;

  float grid_size, grid_density;
  fx_float3 gridSize, gridDelta, gridMin, gridMax;
  fx_int3 gridRes, gridScanMax;
  int gridSrch, gridTotal, gridAdjCnt, gridActive;
  int gridAdj[64];

  float bound_slope, bound_stiff, bound_friction, bound_damp;
  float bound_wall_force, bound_wall_freq;
  fx_float3 bound_min, bound_max;
  fx_float3 init_min, init_max;

  fx_float3 grav_dir, grav_pos, gravity;
  float grav_amt;

  fx_float3 emit_rate, emit_pos, emit_ang, emit_dang, emit_spread;

  fx_int3 brickRes;

  int pemit;
  int pmem;
} FParams_t;

/* Device-side pointer table of one DataX (datax.h:37-47). */
typedef struct cuDataX {
  uint64_t mbuf[FX_DMAXBUF];
} cuDataX;

#ifdef __cplusplus
}
static_assert(sizeof(FParams_t) == 688, "FParams_t must match fluid.h:45-82");
static_assert(sizeof(cuDataX) == 128, "cuDataX must match datax.h:37-47");
#endif

#endif /* FLUIDSX_PARAMS_H */
