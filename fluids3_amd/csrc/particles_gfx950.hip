// particles_gfx950.hip -- the SPH step as hand-written gfx950 (CDNA4, wave64)
// kernels, built into one code object that keeps the reference's module
// contract: the same extern "C" kernel names and argument lists
// (fluids5.0/src/particles.cuh:29-42) and the same four __constant__ symbols
// (fluids5.0/src/particles.cu:17-20), so a host that loads kernels and
// globals by name can drive it.
//
// Arithmetic contract (shared with the CPU oracle, which evaluates the same
// expression trees): the file is compiled with -ffp-contract=off, every fused
// multiply-add is an explicit fmaf(), sqrtf and '/' are IEEE correctly
// rounded.  With the stable in-cell order produced by the counting sort the
// FP32 results are then reproducible run to run and comparable bit for bit.
#include <hip/hip_runtime.h>

#include "fluidsx_params.h"

typedef unsigned int uint;

__constant__ FParams_t FParams;
__constant__ cuDataX FPnts;
__constant__ cuDataX FPntTmp;
__constant__ cuDataX FAccel;

#define SCAN_BLOCKSIZE 512  // threads per scan block; each covers 1024 cells
#define FX_RANK_CAP 2048    // cells with more members keep atomic order

// ---------------------------------------------------------------- helpers

template <class T>
__device__ __forceinline__ T* pbuf(const cuDataX& d, int id) {
  return reinterpret_cast<T*>(d.mbuf[id]);
}

__device__ __forceinline__ fx_float3 ld3(const fx_float3* __restrict__ p, int i) {
  return p[i];
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

// ----------------------------------------------------------- fill (memset)

// Fill used instead of hipMemsetAsync so that a captured step graph holds
// kernel nodes only.
extern "C" __global__ void fxFillUint(uint* __restrict__ p, uint v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ insert

// Cell of each particle and its count slot.  Follows particles.cu:24-54:
//   gc = int((p - gridMin) * gridDelta), gs = (gc.y*res.z + gc.z)*res.x + gc.x,
//   in range iff 1 <= gc <= gridScanMax on every axis, else GRID_UNDEF.
// The range test is done on the float before truncation (trunc(g) >= 1 iff
// g >= 1; trunc(g) <= M iff g < M+1), which is the same predicate for every
// finite input and sends NaN to GRID_UNDEF.  Out-of-range particles are
// counted in slot gridTotal (the tail bucket).
extern "C" __global__ void insertParticles(int pnum) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pnum) return;
  const fx_float3 p = ld3(pbuf<const fx_float3>(FPnts, FPOS), i);
  const float gx = (p.x - FParams.gridMin.x) * FParams.gridDelta.x;
  const float gy = (p.y - FParams.gridMin.y) * FParams.gridDelta.y;
  const float gz = (p.z - FParams.gridMin.z) * FParams.gridDelta.z;
  int* gcell = pbuf<int>(FPnts, FGCELL);
  int* gndx = pbuf<int>(FPnts, FGNDX);
  uint* cnt = pbuf<uint>(FAccel, AGRIDCNT);
  const bool in = gx >= 1.0f && gx < (float)(FParams.gridScanMax.x + 1) &&
                  gy >= 1.0f && gy < (float)(FParams.gridScanMax.y + 1) &&
                  gz >= 1.0f && gz < (float)(FParams.gridScanMax.z + 1);
  int gs;
  if (in) {
    gs = ((int)gy * FParams.gridRes.z + (int)gz) * FParams.gridRes.x + (int)gx;
    gcell[i] = gs;
  } else {
    gs = FParams.gridTotal;
    gcell[i] = FX_GRID_UNDEF;
  }
  gndx[i] = (int)atomicAdd(&cnt[gs], 1u);
}

// ------------------------------------------------------------- prefix scan

// Block-local scan of 1024 uints, reference semantics (particles.cu:333-369):
// block b covers [1024b, 1024b+1024); zeroff=1 gives the exclusive scan,
// zeroff=0 the inclusive one; aux[b] (if given) receives the block total.
// Each thread owns two adjacent elements; waves scan with DPP, the eight wave
// totals are combined through LDS.
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

// ------------------------------------------------------------ counting sort

// Index sort: AGRID[off[cell] + gndx] = i, the member list of every cell in
// insertion (atomic) order.  Same role as fluids3.2's countingSortIndex
// (fluids3.2/fluids/fluid_system_kern.cu:139-150); reads the pre-sort set
// FPntTmp.  Out-of-range particles fill the tail bucket.
extern "C" __global__ void countingSortIndex(int pnum) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pnum) return;
  const int c = pbuf<const int>(FPntTmp, FGCELL)[i];
  const int cell = (c == FX_GRID_UNDEF) ? FParams.gridTotal : c;
  const uint* off = pbuf<const uint>(FAccel, AGRIDOFF);
  pbuf<uint>(FAccel, AGRID)[off[cell] + pbuf<const int>(FPntTmp, FGNDX)[i]] = (uint)i;
}

// Full counting sort FPntTmp -> FPnts (particles.cu:64-95) with a stable
// in-cell order: a particle's slot in its cell is the number of members with
// a smaller pre-sort index (read from the member list), so the sorted order
// equals a stable sort by cell key and does not depend on atomic timing.
// Cells above FX_RANK_CAP members keep the atomic order.  The sort moves the
// state that outlives a step (pos, vel, veval, clr, id) and the fresh gcell
// and gndx; press and force are recomputed by the next two stages.
extern "C" __global__ void countingSortFull(int pnum) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pnum) return;
  const int c = pbuf<const int>(FPntTmp, FGCELL)[i];
  const int cell = (c == FX_GRID_UNDEF) ? FParams.gridTotal : c;
  const uint* off = pbuf<const uint>(FAccel, AGRIDOFF);
  const uint* cnt = pbuf<const uint>(FAccel, AGRIDCNT);
  const uint* grid = pbuf<const uint>(FAccel, AGRID);
  const uint o = off[cell];
  const uint n = cnt[cell];
  int rank;
  if (n <= FX_RANK_CAP) {
    rank = 0;
    for (uint k = 0; k < n; k++) rank += (grid[o + k] < (uint)i) ? 1 : 0;
  } else {
    rank = pbuf<const int>(FPntTmp, FGNDX)[i];
  }
  const uint s = o + (uint)rank;
  pbuf<fx_float3>(FPnts, FPOS)[s] = pbuf<const fx_float3>(FPntTmp, FPOS)[i];
  pbuf<fx_float3>(FPnts, FVEL)[s] = pbuf<const fx_float3>(FPntTmp, FVEL)[i];
  pbuf<fx_float3>(FPnts, FVEVAL)[s] = pbuf<const fx_float3>(FPntTmp, FVEVAL)[i];
  pbuf<uint>(FPnts, FCLR)[s] = pbuf<const uint>(FPntTmp, FCLR)[i];
  pbuf<uint>(FPnts, FID)[s] = pbuf<const uint>(FPntTmp, FID)[i];
  pbuf<int>(FPnts, FGCELL)[s] = c;
  pbuf<int>(FPnts, FGNDX)[s] = rank;
}

// ----------------------------------------------------------------- density

// Row decomposition of the stencil: gridAdj[(y*srch + z)*srch + x] =
// (y*res.z + z)*res.x + x (particles.cpp:614-618) is contiguous in x, so the
// srch cells of one (y, z) row are one contiguous run of sorted particles,
// visited in the same order as the reference's per-cell loop.
__device__ __forceinline__ void row_range(int base, int r, int srch, const uint* __restrict__ off,
                                          const uint* __restrict__ cnt, int& jb, int& je) {
  const int c0 = base + FParams.gridAdj[r * srch];
  const int c1 = c0 + srch - 1;
  jb = (int)off[c0];
  je = (int)(off[c1] + cnt[c1]);
}

// Density (particles.cu:120-157): rho_i = m * poly6 * sum_{0<d2<rd2}
// ((rd2 - d2) * d2s)^3 with d2 in world units, stored in FPRESS; an isolated
// particle gets 1.0.
extern "C" __global__ void computePressure(int pnum) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pnum) return;
  const int gc = pbuf<const int>(FPnts, FGCELL)[i];
  if (gc == FX_GRID_UNDEF) return;
  const fx_float3* __restrict__ pos = pbuf<const fx_float3>(FPnts, FPOS);
  const uint* __restrict__ off = pbuf<const uint>(FAccel, AGRIDOFF);
  const uint* __restrict__ cnt = pbuf<const uint>(FAccel, AGRIDCNT);
  const fx_float3 p = pos[i];
  const int base = gc - (FParams.gridRes.z + 1) * FParams.gridRes.x - 1;
  const int srch = FParams.gridSrch;
  const float rd2 = FParams.rd2, d2 = FParams.d2;
  float sum = 0.0f;
  for (int r = 0; r < srch * srch; r++) {
    int jb, je;
    row_range(base, r, srch, off, cnt, jb, je);
    for (int j = jb; j < je; j++) {
      const fx_float3 q = pos[j];
      const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
      const float dsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
      if (dsq < rd2 && dsq > 0.0f) {
        const float t = (rd2 - dsq) * d2;
        sum = fmaf(t * t, t, sum);
      }
    }
  }
  sum = sum * FParams.pmass * FParams.poly6kern;
  if (sum == 0.0f) sum = 1.0f;
  pbuf<float>(FPnts, FPRESS)[i] = sum;
}

// ------------------------------------------------------------------- force

// Force (particles.cu:188-225): for 0 < d2 < rd2, r = sqrt(d2 * d2s),
// c = h - r, p = (rho - rho0) * k,
// pterm = (ss * -0.5) * c * spiky * (pi + pj) / r,
// F += (pterm * dist + vterm * (veval_j - veval_i)) * c * (1 / (rho_i rho_j)).
extern "C" __global__ void computeForce(int pnum) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pnum) return;
  const int gc = pbuf<const int>(FPnts, FGCELL)[i];
  if (gc == FX_GRID_UNDEF) return;
  const fx_float3* __restrict__ pos = pbuf<const fx_float3>(FPnts, FPOS);
  const fx_float3* __restrict__ vev = pbuf<const fx_float3>(FPnts, FVEVAL);
  const float* __restrict__ rho = pbuf<const float>(FPnts, FPRESS);
  const uint* __restrict__ off = pbuf<const uint>(FAccel, AGRIDOFF);
  const uint* __restrict__ cnt = pbuf<const uint>(FAccel, AGRIDCNT);
  const fx_float3 p = pos[i];
  const fx_float3 vi = vev[i];
  const float di = rho[i];
  const float rest = FParams.prest_dens, stiff = FParams.pintstiff;
  const float pi = (di - rest) * stiff;
  const float h = FParams.psmoothradius, rd2 = FParams.rd2, d2 = FParams.d2;
  const float ssh = FParams.sim_scale * -0.5f;
  const float spiky = FParams.spikykern, vterm = FParams.vterm;
  const int base = gc - (FParams.gridRes.z + 1) * FParams.gridRes.x - 1;
  const int srch = FParams.gridSrch;
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  for (int r = 0; r < srch * srch; r++) {
    int jb, je;
    row_range(base, r, srch, off, cnt, jb, je);
    for (int j = jb; j < je; j++) {
      const fx_float3 q = pos[j];
      const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
      const float dsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
      if (dsq < rd2 && dsq > 0.0f) {
        const float rr = sqrtf(dsq * d2);
        const float c = h - rr;
        const float dj = rho[j];
        const float pj = (dj - rest) * stiff;
        const float pterm = ssh * c * spiky * (pi + pj) / rr;
        const float inv = 1.0f / (di * dj);
        const fx_float3 vj = vev[j];
        const float sx = fmaf(pterm, dx, vterm * (vj.x - vi.x));
        const float sy = fmaf(pterm, dy, vterm * (vj.y - vi.y));
        const float sz = fmaf(pterm, dz, vterm * (vj.z - vi.z));
        fx = fmaf(sx * c, inv, fx);
        fy = fmaf(sy * c, inv, fy);
        fz = fmaf(sz * c, inv, fz);
      }
    }
  }
  pbuf<fx_float3>(FPnts, FFORCE)[i] = fx_float3{fx, fy, fz};
}

// ------------------------------------------------------ neighbour statistics

// Sums over all in-range particles of the stencil candidates (members of the
// stencil cells, self included) and of the in-radius pairs (0 < d2 < rd2),
// on the cell-sorted set: out[0] += C, out[1] += P.  The measurement behind
// the flop model of the density and force stages; the role of the
// reference's neighbour-count validation (fluids4.0/fluid_system.cpp:959-1011).
extern "C" __global__ void fxCountPairs(unsigned long long* out, int pnum) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long nc = 0, np = 0;
  if (i < pnum) {
    const int gc = pbuf<const int>(FPnts, FGCELL)[i];
    if (gc != FX_GRID_UNDEF) {
      const fx_float3* __restrict__ pos = pbuf<const fx_float3>(FPnts, FPOS);
      const uint* __restrict__ off = pbuf<const uint>(FAccel, AGRIDOFF);
      const uint* __restrict__ cnt = pbuf<const uint>(FAccel, AGRIDCNT);
      const fx_float3 p = pos[i];
      const int base = gc - (FParams.gridRes.z + 1) * FParams.gridRes.x - 1;
      const int srch = FParams.gridSrch;
      for (int r = 0; r < srch * srch; r++) {
        int jb, je;
        row_range(base, r, srch, off, cnt, jb, je);
        nc += (unsigned long long)(je - jb);
        for (int j = jb; j < je; j++) {
          const fx_float3 q = pos[j];
          const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
          const float dsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
          np += (dsq < FParams.rd2 && dsq > 0.0f) ? 1u : 0u;
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

// ----------------------------------------------------------------- advance

// Integrator (particles.cu:237-322): penalty walls on bound_min/bound_max
// (Y-min with slope and friction, Y-max with friction, X-min moving wall whose
// stiffness term is scaled by 'wall', X-max, Z-min, Z-max), gravity, the
// acceleration limit, the velocity limit applied to the old velocity, and the
// leapfrog update.  GRID_UNDEF particles are parked at (-1000)^3 with zero
// velocity.
__device__ __forceinline__ void advance_one(int i, float time, float dt, float ss) {
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
  const float m = FParams.pmass;
  float ax = f.x * m, ay = f.y * m, az = f.z * m;
  const float rad = FParams.pradius, stiff = FParams.bound_stiff, damp = FParams.bound_damp;
  const float slope = FParams.bound_slope, fric = FParams.bound_friction;
  const fx_float3 bmin = FParams.bound_min, bmax = FParams.bound_max;
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
  const float sarg = time * FParams.bound_wall_freq + pos.z / 400.0f;
  const float wall = (float)(((double)sinf(sarg) * 0.5 + 0.5) * (double)FParams.bound_wall_force);
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
  ax += FParams.gravity.x;
  ay += FParams.gravity.y;
  az += FParams.gravity.z;
  // acceleration limit
  float speed = ax * ax + ay * ay + az * az;
  if (speed > FParams.AL2) {
    const float s = FParams.AL / sqrtf(speed);
    ax *= s; ay *= s; az *= s;
  }
  // velocity limit, applied to the old velocity
  speed = vel.x * vel.x + vel.y * vel.y + vel.z * vel.z;
  if (speed > FParams.VL2) {
    const float s = FParams.VL / sqrtf(speed);
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
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numPnts) return;
  advance_one(i, time, dt, ss);
}

// Same integrator with the simulation time read from device memory, for
// captured step graphs whose launches cannot carry a per-step time argument.
// fxClockTick advances that clock by dt (the host's AdvanceTime,
// particles.cpp:396-400).
extern "C" __global__ void advanceParticlesClock(const float* clock, float dt, float ss,
                                                 int numPnts) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numPnts) return;
  advance_one(i, clock[0], dt, ss);
}

extern "C" __global__ void fxClockTick(float* clock, float dt) {
  if (blockIdx.x == 0 && threadIdx.x == 0) clock[0] = clock[0] + dt;
}
