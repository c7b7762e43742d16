// adjoint_gfx950.hip -- reverse-mode (adjoint) kernels of the SPH step for
// gfx950 (CDNA4, wave64).  No reference counterpart: the reference has no
// differentiable path.  They reverse computePressure, computeForce and
// advanceParticles (fluids5.0/src/particles.cu:120-157, 188-225, 237-322) as
// implemented in particles_gfx950.hip, over a recorded step: the step's
// cell-sorted positions, velocities, evaluation velocities, forces,
// densities, ids and cells and its count/offset tables.  Neighbour sets,
// cells, branch masks and the sort permutation are taken from that forward
// step as piecewise-constant data.
//
// Adjoint variables of the state live in the slot order of the state they
// belong to: those of step t's outputs in step t's sorted order, those of
// its inputs in the order before its sort, which the step's recorded sort
// permutation src (slot s holds the particle that was at src[s]) gives.
// Each step reads one set and writes the other, and the two alternate.
// Per step, for the step's sorted slots:
//   adjAdvance  (gx', gv', ge') of the outputs -> gF, partial gx/ge of the
//               inputs (sorted frame), gv of the velocity input (written at
//               src[s], the order before the sort);
//   adjForce    gather over in-radius neighbours j (dl = gF_i - gF_j,
//               u = x_i - x_j, dve = ve_j - ve_i, inv = 1/(r rho_i rho_j)):
//                 gx_i   += A dl + (d2/r) u (A_r (u.dl) + B_r (dve.dl))
//                 ge_i   -= B dl
//                 grho_i += kp c^2 inv k (u.dl) - (A (u.dl) + B (dve.dl)) / rho_i
//               and the per-particle terms of dL/dk and dL/dvisc;
//   adjDensity  gx_i += sum_j (grho_i + grho_j) (-6 m poly6 d2) t^2 u, then
//               the step's input adjoints go to src[s].
// Every pair term is gathered from i's side (A and B are symmetric in i, j),
// so there are no atomics and results are reproducible bit for bit.  FP64 is
// used only for the loss and the two scalar-parameter sums: per-block FP64
// partials, reduced per scene in a fixed order.
//
// Arithmetic contract as in particles_gfx950.hip: -ffp-contract=off, explicit
// fmaf, IEEE sqrt and division; the CPU adjoint oracle evaluates the same
// trees.
#include <hip/hip_runtime.h>

#include <climits>

#include "fluidsx_params.h"
#include "fx_trig.hpp"

typedef unsigned int uint;

#define ADJ_THREADS 64   // one wave per block
#define RED_THREADS 256

template <class T>
__device__ __forceinline__ T* gptr(T* p) {
  typedef __attribute__((address_space(1))) T gT;
  return (T*)(gT*)p;
}

__device__ __forceinline__ const FParams_t* cparams(const FParams_t* p) {
  typedef const __attribute__((address_space(4))) FParams_t cP;
  return (const FParams_t*)(cP*)p;
}

// Everything one adjoint launch reads or writes.  Particle vectors are fx_float3
// arrays; "sorted" arrays are indexed by the step's slot, "id" arrays by
// particle id.
struct AdjArgs {
  const FParams_t* params;  // one per scene
  const int* pbase;         // first particle of each scene (nscene + 1)
  const int2* btab;         // (scene, block inside the scene) per launched block
  // recorded forward step (sorted frame)
  const fx_float3* pos;
  const fx_float3* vel;
  const fx_float3* veval;
  const fx_float3* force;
  const float* rho;
  const uint* id;  // adjLoss: the final state's ids; the step kernels: the step's sort permutation src
  const int* gcell;
  const uint* cnt;
  const uint* off;
  const fx_float3* target;  // by id (loss only)
  // adjoint state of the step's outputs (its sorted order) ...
  fx_float3* gx;
  fx_float3* gv;
  fx_float3* ge;
  // ... and of its inputs (the order before its sort)
  fx_float3* gxp;
  fx_float3* gvp;
  fx_float3* gep;
  // per-step work (sorted frame)
  fx_float3* gF;
  fx_float3* gxs;
  fx_float3* ges;
  float* grho;
  double* partial;  // 3 doubles per launched block: loss, dL/dk, dL/dvisc
  float4* arec;     // per slot: (pos, rho), (veval, 0), (gF, 0), written by adjAdvance
  float4* drec;     // per slot: (pos, grho), written by adjForce
  const uint* bits; // the recorded step's neighbour bitmask, word w of slot i at [w * nbits + i]
  int nbits;
  float time;
  int accumulate;  // adjForce: add to the block's partials instead of overwriting
  int ncap;        // bitmask words a slot holds; a particle needing more recorded none
};

struct Slot {
  const FParams_t* P;
  int i, p0, p1, scene;
  int blk;  // logical block: entry of the block table and of the partials
};

// XCD-aware block order, as in the forward pair kernels: MI355X hands
// consecutive workgroups to its eight XCDs round-robin, so runs of
// ADJ_XCD_RUN consecutive logical blocks (neighbouring particles, which share
// neighbours) are given to one XCD and the runs rotate over the XCDs.
// Dispatch slot b is the k-th block of XCD x = b % 8 and runs logical block
// (run * 8 + x) * ADJ_XCD_RUN + k % ADJ_XCD_RUN with run = k / ADJ_XCD_RUN;
// blocks past the last full round keep their index.
#ifndef ADJ_XCD_RUN
#define ADJ_XCD_RUN 128
#endif
__device__ __forceinline__ int adj_block() {
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int round = 8 * ADJ_XCD_RUN;
  if (b >= nb / round * round) return b;
  const int x = b % 8, k = b / 8;
  return ((k / ADJ_XCD_RUN) * 8 + x) * ADJ_XCD_RUN + k % ADJ_XCD_RUN;
}

__device__ __forceinline__ Slot slot_of(const AdjArgs& A) {
  Slot s;
  s.blk = adj_block();
  const int2 e = A.btab[s.blk];
  s.P = cparams(A.params + e.x);
  s.scene = e.x;
  s.p0 = A.pbase[e.x];
  s.p1 = A.pbase[e.x + 1];
  s.i = s.p0 + e.y * ADJ_THREADS + (int)threadIdx.x;
  return s;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block sum of one double per thread in a fixed order (wave butterflies, then
// waves in index order); thread 0 gets the result.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  if (NT == 64) return v;
  const int w = (int)threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < NT / 64; k++) s += lds[k];
  __syncthreads();
  return s;
}

// ------------------------------------------------------------------- loss
// L_s = (1/n_s) sum |x - x*|^2 over scene s (FP64), seed gx = (2/n_s)(x - x*),
// gv = ge = 0, in the final state's slot order.  pos/id: the final state.
extern "C" __global__ void __launch_bounds__(ADJ_THREADS) adjLoss(AdjArgs A) {
  __shared__ double lds[1];
  const Slot S = slot_of(A);
  double l = 0.0;
  if (S.i < S.p1) {
    const uint id = A.id[S.i];
    const fx_float3 x = A.pos[S.i], t = A.target[id];
    const float scale = 2.0f / (float)(S.p1 - S.p0);
    const float dx = x.x - t.x, dy = x.y - t.y, dz = x.z - t.z;
    l = (double)dx * (double)dx + (double)dy * (double)dy + (double)dz * (double)dz;
    A.gx[S.i] = fx_float3{dx * scale, dy * scale, dz * scale};
    A.gv[S.i] = fx_float3{0.0f, 0.0f, 0.0f};
    A.ge[S.i] = fx_float3{0.0f, 0.0f, 0.0f};
  }
  l = block_sum<ADJ_THREADS>(l, lds);
  if (threadIdx.x == 0) A.partial[3 * S.blk] = l;
}

// ------------------------------------------------------------ rollout frame
// For the length of a rollout the particles are numbered by their slot at its
// start (FID := slot): the adjoint of the rollout's first step comes out in
// that order, and the loss reads the target through ids that stay close to
// the final state's slot order (particles move little between steps).
// adjFrameIn saves the original ids, renumbers and moves the target into the
// frame; adjFrameOut restores the ids of the current state; adjUnframe moves
// the three gradients back to original-id order.
extern "C" __global__ void adjFrameIn(uint* fid, uint* id0, const fx_float3* target, fx_float3* target_f, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint id = fid[s];
  id0[s] = id;
  fid[s] = (uint)s;
  target_f[s] = target[id];
}

extern "C" __global__ void adjFrameOut(uint* fid, const uint* id0, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) fid[s] = id0[fid[s]];
}

extern "C" __global__ void adjUnframe(const fx_float3* g, fx_float3* out, const uint* id0, int n) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const size_t id = id0[a];
  for (int v = 0; v < 3; v++) out[(size_t)v * n + id] = g[(size_t)v * n + a];
}

// ------------------------------------------------------------------- tape
// One recorded step's tape slot, written in one pass: pos, vel, veval, force
// (12 B each), rho, id, gcell (4 B each) of the n particles, then the cell
// counts and offsets of the cells-entry table, each array contiguous in that
// order.  Thread a copies particle a and table entry a.
struct TapeSrc {
  const fx_float3* v3[4];  // pos, vel, veval, force
  const uint* s1[3];       // rho (as bits), id, gcell
  const uint* cnt;
  const uint* off;
};

extern "C" __global__ void adjTapeRecord(TapeSrc S, char* dst, int n, int cells) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t N = (size_t)n;
  if (a < n) {
    // as twelve floats and three words (an array of fx_float3 copies would
    // be placed in LDS by the compiler, 48 KB per block)
    float v[12];
    uint w[3];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const float* f = reinterpret_cast<const float*>(S.v3[k]) + 3 * (size_t)a;
      v[3 * k] = f[0];
      v[3 * k + 1] = f[1];
      v[3 * k + 2] = f[2];
    }
#pragma unroll
    for (int k = 0; k < 3; k++) w[k] = S.s1[k][a];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      float* f = reinterpret_cast<float*>(dst + 12 * N * k) + 3 * (size_t)a;
      f[0] = v[3 * k];
      f[1] = v[3 * k + 1];
      f[2] = v[3 * k + 2];
    }
#pragma unroll
    for (int k = 0; k < 3; k++) reinterpret_cast<uint*>(dst + 48 * N + 4 * N * k)[a] = w[k];
  }
  if (a < cells) {
    const uint c = S.cnt[a], o = S.off[a];
    uint* t = reinterpret_cast<uint*>(dst + 60 * N);
    t[a] = c;
    t[(size_t)cells + a] = o;
  }
}

// A checkpoint: pos, vel, veval (12 B each), clr, id (4 B each) of the n
// particles, contiguous in that order in ckpt.  save != 0 copies the state
// into the checkpoint, save == 0 back into the particle buffers.
struct CkptPtr {
  fx_float3* v3[3];  // pos, vel, veval
  uint* s1[2];       // clr, id
};

extern "C" __global__ void adjCheckpoint(CkptPtr S, char* ckpt, int n, int save) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const size_t N = (size_t)n;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    fx_float3* c = reinterpret_cast<fx_float3*>(ckpt + 12 * N * k) + a;
    if (save)
      *c = S.v3[k][a];
    else
      S.v3[k][a] = *c;
  }
#pragma unroll
  for (int k = 0; k < 2; k++) {
    uint* c = reinterpret_cast<uint*>(ckpt + 36 * N + 4 * N * k) + a;
    if (save)
      *c = S.s1[k][a];
    else
      S.s1[k][a] = *c;
  }
}

// ---------------------------------------------------------------- advance
extern "C" __global__ void __launch_bounds__(ADJ_THREADS) adjAdvance(AdjArgs A) {
  const Slot S = slot_of(A);
  if (S.i >= S.p1) return;
  const FParams_t& P = *S.P;
  const int s = S.i;
  const uint src = A.id[s];
  const fx_float3 gxo = A.gx[s], gvo = A.gv[s], geo = A.ge[s];
  if (A.gcell[s] == FX_GRID_UNDEF) {  // parked: pos and vel become constants, veval passes through
    const fx_float3 z = {0.0f, 0.0f, 0.0f};
    A.gxs[s] = z;
    A.ges[s] = geo;
    A.gvp[src] = z;
    return;
  }
  const fx_float3 pos = A.pos[s], vel = A.vel[s], ve = A.veval[s], f = A.force[s];
  const float m = P.pmass, dt = P.dt, ss = P.sim_scale;
  float ax = f.x * m, ay = f.y * m, az = f.z * m;
  const float rad = P.pradius, stiff = P.bound_stiff, damp = P.bound_damp;
  const float slope = P.bound_slope, fric = P.bound_friction;
  const fx_float3 bmin = P.bound_min, bmax = P.bound_max;
  float diff, adj, nx, ny, nz;
  // forward replay (advance_one of particles_gfx950.hip): pre-limit
  // acceleration and the active walls
  diff = rad - (pos.y - (bmin.y + (pos.x - bmin.x) * slope)) * ss;
  const bool w1 = diff > FX_EPSILON;
  if (w1) {
    nx = -slope; ny = 1.0f - slope; nz = 0.0f;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    nx *= adj; ny *= adj; nz *= adj;
    ax += nx - ve.x * fric; ay += ny - ve.y * fric; az += nz - ve.z * fric;
  }
  diff = rad - (bmax.y - pos.y) * ss;
  const bool w2 = diff > FX_EPSILON;
  if (w2) {
    nx = 0.0f; ny = -1.0f; nz = 0.0f;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    nx *= adj; ny *= adj; nz *= adj;
    ax += nx - ve.x * fric; ay += ny - ve.y * fric; az += nz - ve.z * fric;
  }
  const float sarg = A.time * P.bound_wall_freq + pos.z / 400.0f;
  const float wall = (float)(((double)fx_sinf(sarg) * 0.5 + 0.5) * (double)P.bound_wall_force);
  diff = rad - (pos.x - (bmin.x + wall)) * ss;
  const float diff3 = diff;
  const bool w3 = diff > FX_EPSILON;
  if (w3) {
    nx = 1.0f; ny = 0.0f; nz = 0.0f;
    adj = wall * stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    ax += nx * adj; ay += ny * adj; az += nz * adj;
  }
  diff = rad - (bmax.x - pos.x) * ss;
  const bool w4 = diff > FX_EPSILON;
  if (w4) {
    nx = -1.0f; ny = 0.0f; nz = 0.0f;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    ax += nx * adj; ay += ny * adj; az += nz * adj;
  }
  diff = rad - (pos.z - bmin.z) * ss;
  const bool w5 = diff > FX_EPSILON;
  if (w5) {
    nx = 0.0f; ny = 0.0f; nz = 1.0f;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    ax += nx * adj; ay += ny * adj; az += nz * adj;
  }
  diff = rad - (bmax.z - pos.z) * ss;
  const bool w6 = diff > FX_EPSILON;
  if (w6) {
    nx = 0.0f; ny = 0.0f; nz = -1.0f;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    ax += nx * adj; ay += ny * adj; az += nz * adj;
  }
  ax += P.gravity.x;
  ay += P.gravity.y;
  az += P.gravity.z;
  const float sa2 = ax * ax + ay * ay + az * az;
  const float sv2 = vel.x * vel.x + vel.y * vel.y + vel.z * vel.z;

  // reverse of x' = x + vn dt/ss, v' = vn, ve' = (v_l + vn)/2, vn = a_l dt + v_l
  const float dts = dt / ss;
  const float gnx = gxo.x * dts + gvo.x + geo.x * 0.5f;
  const float gny = gxo.y * dts + gvo.y + geo.y * 0.5f;
  const float gnz = gxo.z * dts + gvo.z + geo.z * 0.5f;
  float gvx = gnx + geo.x * 0.5f, gvy = gny + geo.y * 0.5f, gvz = gnz + geo.z * 0.5f;
  float gax = gnx * dt, gay = gny * dt, gaz = gnz * dt;
  if (sv2 > P.VL2) {  // v_l = v VL/|v|: J = (VL/|v|)(I - v v^T/|v|^2)
    const float sv = P.VL / sqrtf(sv2);
    const float d = (vel.x * gvx + vel.y * gvy + vel.z * gvz) / sv2;
    gvx = (gvx - vel.x * d) * sv;
    gvy = (gvy - vel.y * d) * sv;
    gvz = (gvz - vel.z * d) * sv;
  }
  if (sa2 > P.AL2) {
    const float sa = P.AL / sqrtf(sa2);
    const float d = (ax * gax + ay * gay + az * gaz) / sa2;
    gax = (gax - ax * d) * sa;
    gay = (gay - ay * d) * sa;
    gaz = (gaz - az * d) * sa;
  }
  A.gvp[src] = fx_float3{gvx, gvy, gvz};
  A.arec[3 * (size_t)s] = make_float4(pos.x, pos.y, pos.z, A.rho[s]);
  A.arec[3 * (size_t)s + 1] = make_float4(ve.x, ve.y, ve.z, 0.0f);
  A.arec[3 * (size_t)s + 2] = make_float4(gax * m, gay * m, gaz * m, 0.0f);
  float px = gxo.x, py = gxo.y, pz = gxo.z;
  float ex = 0.0f, ey = 0.0f, ez = 0.0f;
  if (w1) {
    nx = -slope; ny = 1.0f - slope;
    const float dn = nx * gax + ny * gay;
    const float t = stiff * dn * ss;
    px += t * slope;
    py -= t;
    const float dd = damp * dn;
    ex -= dd * nx + fric * gax;
    ey -= dd * ny + fric * gay;
    ez -= fric * gaz;
  }
  if (w2) {
    const float dn = -gay;
    py += stiff * dn * ss;
    ex -= fric * gax;
    ey -= fric * gay - damp * dn;
    ez -= fric * gaz;
  }
  if (w3) {
    px -= gax * (wall * stiff * ss);
    if (P.bound_wall_force != 0.0f) {
      const float dwall = fx_cosf(sarg) * (0.5f / 400.0f) * P.bound_wall_force;
      pz += gax * (stiff * dwall * (diff3 + wall * ss));
    }
    ex -= damp * gax;
  }
  if (w4) {
    px -= gax * (stiff * ss);
    ex -= damp * gax;
  }
  if (w5) {
    pz -= gaz * (stiff * ss);
    ez -= damp * gaz;
  }
  if (w6) {
    pz -= gaz * (stiff * ss);
    ez -= damp * gaz;
  }
  A.gxs[s] = fx_float3{px, py, pz};
  A.ges[s] = fx_float3{ex, ey, ez};
}

// ------------------------------------------------------- neighbour rings
//
// The recorded forward step's density kernel left, per particle, one bit per
// stencil candidate (set for the in-radius pairs; row r occupies a
// word-aligned run of ceil(len/32) words at [w * N + i]).  Both pair kernels
// below turn those bits into neighbour indices in a per-lane ring in LDS and
// evaluate queued pairs in wave-wide batches (every lane takes as many pairs
// as the shortest active queue holds), so the pair arithmetic runs with
// nearly all lanes busy and no candidate is tested twice.  Each lane visits
// its pairs in candidate order (the order of the CPU oracle's loops).  One
// wave per block; a particle whose candidates needed more than ncap
// words has no bits and tests its candidates in place.
#define AQ_SLOTS 1024  // ring slots per block: 16 per lane

__device__ __forceinline__ int wave_max_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_min_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}

// Row r of the stencil at base: candidates [jb, je).  A row ends where the
// cell after its last one begins (an in-range particle's row is followed by
// a cell of the same grid row), so two loads, not three.
__device__ __forceinline__ void row_of(const FParams_t& P, int base, int r, const uint* off, const uint* cnt,
                                       int& jb, int& je) {
  const int srch = P.gridSrch;
  const int c0 = base + P.gridAdj[r * srch];
  jb = (int)off[c0];
  je = (int)off[c0 + srch];
}

__device__ __forceinline__ int words_of(const FParams_t& P, int base, const uint* off, const uint* cnt) {
  int w = 0;
  for (int r = 0; r < P.gridSrch * P.gridSrch; r++) {
    int jb, je;
    row_of(P, base, r, off, cnt, jb, je);
    w += (je - jb + 31) >> 5;
  }
  return w;
}

// Walks the lane's recorded pairs through the ring; op.eval(j, ok) evaluates
// pair j and keeps the result only when ok (a false call is a padding slot:
// j = 0, loads stay in bounds, nothing is accumulated).  All 64 lanes must
// call it (wave-wide maxima).
template <class Op>
__device__ __forceinline__ void ring_walk(const FParams_t& P, const AdjArgs& A, int i, int base, bool use,
                                          uint* fq, Op& op) {
  const int t = threadIdx.x;
  const int fqn = AQ_SLOTS / 64;
  const int srch = P.gridSrch;
  const size_t NN = (size_t)A.nbits;
  int head = 0, qn = 0, W = 0;
  auto drain = [&](int f) {
#pragma unroll 2
    for (int e = 0; e < f; e++) {
      const bool ok = e < qn;
      const uint j = ok ? fq[head * 64 + t] : 0u;
      op.eval(j, ok);
      if (ok) head = (head + 1 == fqn) ? 0 : head + 1;
    }
    qn = max(qn - f, 0);
  };
  for (int r = 0; r < srch * srch; r++) {
    int jb = 0, je = 0;
    if (use) row_of(P, base, r, A.off, A.cnt, jb, je);
    const int nw = (je - jb + 31) >> 5;
    const int maxw = min(wave_max_i(nw), A.ncap);
    for (int w = 0; w < maxw; w++) {
      uint word = 0;
      if (w < nw) word = A.bits[(size_t)(W + w) * NN + i];
      for (int h = 0; h < 32; h += 16) {
        uint m = (word >> h) & 0xffffu;
        const int pc = __popc(m);
        if (__any(qn + pc > fqn)) {
          const int need = wave_max_i(qn + pc - fqn);
          drain(min(max(need, wave_min_i(use ? qn : 0x7fffffff)), fqn));
        }
        int tail = head + qn;
        if (tail >= fqn) tail -= fqn;
        while (m != 0u) {
          const int b = __builtin_ctz(m);
          m &= m - 1u;
          fq[tail * 64 + t] = (uint)(jb + (w << 5) + h + b);
          tail = (tail + 1 == fqn) ? 0 : tail + 1;
          qn++;
        }
      }
    }
    W += nw;
  }
  drain(min(wave_max_i(qn), fqn));
}

// Walks the lane's recorded pairs through a word queue: each lane reads its
// bitmask words as one stream (row r's words follow row r-1's) and queues
// every non-zero word whole, with the record index of its bit 0 -- two words
// in registers, up to AW_QWORDS more in LDS; the wave then takes the same
// number of pairs from every lane (as many as the shortest queue holds, more
// only when a queue must make room), AW_STEP per step, each lane's in candidate
// order.  Same contract as ring_walk; for stencil rows of SR cells, K LDS
// words per lane (8 K bytes per block).
#ifndef AW_QF
#define AW_QF 8   // LDS words per lane in the force adjoint
#endif
#ifndef AW_QD
#define AW_QD 8   // and in the density adjoint
#endif
#ifndef AW_STEP
#define AW_STEP 3  // pairs taken per drain step
#endif
#ifndef AW_FORCE_ATTR
#define AW_FORCE_ATTR
#endif

template <int SR, int K, class Op>
__device__ __forceinline__ bool queue_walk(const FParams_t& P, const AdjArgs& A, int i, int base, bool use,
                                           uint* lq, Op& op) {
  constexpr int R = SR * SR;
  const int t = threadIdx.x;
  const size_t NN = (size_t)A.nbits;
  // row r covers words [W_r, W_{r+1}) of the lane's bitmask; E_[r] = W_{r+1},
  // and the record of word f, bit b is G[r] + 32 f + b
  int E_[R], G[R];
  int nwords = 0;
#pragma unroll
  for (int r = 0; r < R; r++) {
    int jb = 0, je = 0;
    if (use) row_of(P, base, r, A.off, A.cnt, jb, je);
    G[r] = jb - (nwords << 5);
    nwords += (je - jb + 31) >> 5;
    E_[r] = nwords;
  }
  // a lane whose rows need more words than a tape slot holds recorded none
  // of them: it walks nothing here
  const bool over = use && nwords > A.ncap;
  if (over) {
    nwords = 0;
#pragma unroll
    for (int r = 0; r < R; r++) E_[r] = G[r] = 0;
  }
  use = use && !over;
  uint* Qw = lq;
  uint* Qb = lq + K * 64;
  uint cur = 0u, nxt = 0u;
  int curb = 0, nxtb = 0, qh = 0, qc = 0, qn = 0;
  auto take = [&](uint& j) -> bool {
    const bool ok = cur != 0u;
    j = ok ? (uint)(curb + __builtin_ctz(cur)) : 0u;
    cur &= cur - 1u;
    if (ok && cur == 0u) {
      cur = nxt;
      curb = nxtb;
      nxt = 0u;
      if (qc > 0) {
        nxt = Qw[qh * 64 + t];
        nxtb = (int)Qb[qh * 64 + t];
        qh = (qh + 1 == K) ? 0 : qh + 1;
        qc--;
      }
    }
    return ok;
  };
  auto drain = [&](int f) {
    int e = 0;
    for (; e + AW_STEP <= f; e += AW_STEP) {
      uint j[AW_STEP];
      bool ok[AW_STEP];
      typename Op::Rec rc[AW_STEP];
#pragma unroll
      for (int u = 0; u < AW_STEP; u++) ok[u] = take(j[u]);
#pragma unroll
      for (int u = 0; u < AW_STEP; u++) rc[u] = op.load(j[u]);  // every gather of the step first
#pragma unroll
      for (int u = 0; u < AW_STEP; u++) op.apply(rc[u], ok[u]);
    }
    for (; e < f; e++) {
      uint j;
      const bool ok = take(j);
      op.eval(j, ok);
    }
    qn = max(qn - f, 0);
  };
  const int maxf = wave_max_i(nwords);
  for (int f = 0; f < maxf; f++) {
    const uint word = f < nwords ? A.bits[(size_t)f * NN + i] : 0u;
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
      const int need = wave_max_i(full ? __popc(cur) : 0);
      drain(max(need, wave_min_i(use ? qn : 0x7fffffff)));
    }
    if (word != 0u) {
      const int wb = G[0] + (f << 5);
      if (cur == 0u) {
        cur = word;
        curb = wb;
      } else if (nxt == 0u) {
        nxt = word;
        nxtb = wb;
      } else {
        int tl = qh + qc;
        if (tl >= K) tl -= K;
        Qw[tl * 64 + t] = word;
        Qb[tl * 64 + t] = (uint)wb;
        qc++;
      }
      qn += __popc(word);
    }
  }
  drain(wave_max_i(qn));
  return over;
}

// Stencil rows of 2 or 3 cells through the word queue, 4 through the ring.
// Returns whether the lane's rows need more bitmask words than a tape slot
// holds (then it recorded no bits, and nothing was walked for it).
template <int K, class Op>
__device__ __forceinline__ bool pair_walk(const FParams_t& P, const AdjArgs& A, int i, int base, bool use,
                                          uint* lq, Op& op) {
  if (P.gridSrch == 3) return queue_walk<3, K>(P, A, i, base, use, lq, op);
  if (P.gridSrch == 2) return queue_walk<2, K>(P, A, i, base, use, lq, op);
  const bool over = use && words_of(P, base, A.off, A.cnt) > A.ncap;
  ring_walk(P, A, i, base, use && !over, lq, op);
  return over;
}

// ------------------------------------------------------------------ force

// Pair term of the force adjoint; same expression tree as the CPU oracle.
struct ForceAdj {
  const FParams_t* P;
  const float4* rec;  // (pos, rho), (veval, 0), (gF, 0) per slot
  float px, py, pz, vx, vy, vz, gx, gy, gz, di, pi, idi, qi;
  float kp, stiff, vterm, lap, h, rest, d2;
  float gpx, gpy, gpz, gex, gey, gez, gr, ak, av;

  // the pair's records (load), then its terms added in order (apply)
  struct Rec {
    float4 a, v, q;
  };
  __device__ __forceinline__ Rec load(uint j) const {
    return Rec{rec[3 * (size_t)j], rec[3 * (size_t)j + 1], rec[3 * (size_t)j + 2]};
  }
  __device__ __forceinline__ void eval(uint j, bool ok) { apply(load(j), ok); }
  __device__ __forceinline__ void apply(const Rec& r, bool ok) {
    const float4 a = r.a, v = r.v, q = r.q;
    const float dx = px - a.x, dy = py - a.y, dz = pz - a.z;
    const float dsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const float rr = sqrtf(dsq * d2);
    const float c = h - rr;
    const float dj = a.w;
    const float pj = (dj - rest) * stiff;
    const float inv = 1.0f / (rr * (di * dj));
    const float invd = inv * rr;
    const float ir = inv * (di * dj);
    const float kc2i = (kp * (c * c)) * inv;
    const float pp = pi + pj;
    const float Ac = kc2i * pp;
    const float Bc = (vterm * c) * invd;
    const float lx = gx - q.x, ly = gy - q.y, lz = gz - q.z;
    const float wx = v.x - vx, wy = v.y - vy, wz = v.z - vz;
    const float su = fmaf(dz, lz, fmaf(dy, ly, dx * lx));
    const float sv = fmaf(wz, lz, fmaf(wy, ly, wx * lx));
    const float Ar = -(kp * pp) * (c * (rr + rr + c)) * (invd * (ir * ir));
    const float Br = -vterm * invd;
    const float g = (d2 * ir) * fmaf(Ar, su, Br * sv);
    const float ui = fmaf(dz, gz, fmaf(dy, gy, dx * gx));
    const float wi = fmaf(wz, gz, fmaf(wy, gy, wx * gx));
    const float n_gpx = fmaf(Ac, lx, fmaf(g, dx, gpx));
    const float n_gpy = fmaf(Ac, ly, fmaf(g, dy, gpy));
    const float n_gpz = fmaf(Ac, lz, fmaf(g, dz, gpz));
    const float n_gex = fmaf(-Bc, lx, gex);
    const float n_gey = fmaf(-Bc, ly, gey);
    const float n_gez = fmaf(-Bc, lz, gez);
    const float n_gr = fmaf(kc2i * stiff, su, gr) - fmaf(Ac, su, Bc * sv) * idi;
    const float n_ak = fmaf(kc2i * (qi + (dj - rest)), ui, ak);
    const float n_av = fmaf((lap * c) * invd, wi, av);
    if (ok) {
      gpx = n_gpx; gpy = n_gpy; gpz = n_gpz;
      gex = n_gex; gey = n_gey; gez = n_gez;
      gr = n_gr; ak = n_ak; av = n_av;
    }
  }
};

extern "C" __global__ void __launch_bounds__(ADJ_THREADS) AW_FORCE_ATTR adjForce(AdjArgs A) {
  __shared__ uint fq[(2 * AW_QF * 64 > AQ_SLOTS) ? 2 * AW_QF * 64 : AQ_SLOTS];
  const Slot S = slot_of(A);
  const FParams_t& P = *S.P;
  const int i = S.i;
  const int gc = i < S.p1 ? A.gcell[i] : FX_GRID_UNDEF;
  const bool act = gc != FX_GRID_UNDEF;
  const int base = gc - (P.gridRes.z + 1) * P.gridRes.x - 1;
  ForceAdj F;
  F.P = S.P;
  F.rec = A.arec;
  F.kp = (P.sim_scale * -0.5f) * P.spikykern;
  F.stiff = P.pintstiff;
  F.vterm = P.vterm;
  F.lap = P.lapkern;
  F.h = P.psmoothradius;
  F.rest = P.prest_dens;
  F.d2 = P.d2;
  F.px = F.py = F.pz = F.vx = F.vy = F.vz = F.gx = F.gy = F.gz = 0.0f;
  F.di = 1.0f;
  F.pi = F.qi = 0.0f;
  F.idi = 1.0f;
  F.gpx = F.gpy = F.gpz = F.gex = F.gey = F.gez = F.gr = F.ak = F.av = 0.0f;
  if (act) {
    const float4 a = A.arec[3 * (size_t)i], v = A.arec[3 * (size_t)i + 1], q = A.arec[3 * (size_t)i + 2];
    F.px = a.x; F.py = a.y; F.pz = a.z;
    F.vx = v.x; F.vy = v.y; F.vz = v.z;
    F.gx = q.x; F.gy = q.y; F.gz = q.z;
    F.di = a.w;
    F.pi = (F.di - F.rest) * F.stiff;
    F.idi = 1.0f / F.di;
    F.qi = F.di - F.rest;
  }
  // a particle without recorded bits tests its candidates in place, in the
  // same candidate order (after the walk, which took none of its pairs)
  const bool ovf = pair_walk<AW_QF>(P, A, i, base, act, fq, F);
  if (ovf) {
    const float rd2 = P.rd2;
    for (int r = 0; r < P.gridSrch * P.gridSrch; r++) {
      int jb, je;
      row_of(P, base, r, A.off, A.cnt, jb, je);
      for (int j = jb; j < je; j++) {
        const float4 a = A.arec[3 * (size_t)j];
        const float dx = F.px - a.x, dy = F.py - a.y, dz = F.pz - a.z;
        const float dsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        if (dsq < rd2 && dsq > 0.0f) F.eval((uint)j, true);
      }
    }
  }
  if (act) {
    const fx_float3 x0 = A.gxs[i], e0 = A.ges[i];
    A.gxs[i] = fx_float3{x0.x + F.gpx, x0.y + F.gpy, x0.z + F.gpz};
    A.ges[i] = fx_float3{e0.x + F.gex, e0.y + F.gey, e0.z + F.gez};
    A.grho[i] = F.gr;
    A.drec[i] = make_float4(F.px, F.py, F.pz, F.gr);
  } else if (i < S.p1) {
    A.grho[i] = 0.0f;
  }
  const double sk = wave_sum((double)F.ak);
  const double sv = wave_sum((double)F.av);
  if (threadIdx.x == 0) {
    double* o = A.partial + 3 * S.blk;
    o[1] = A.accumulate ? o[1] + sk : sk;
    o[2] = A.accumulate ? o[2] + sv : sv;
  }
}

// ---------------------------------------------------------------- density

struct DensityAdj {
  const float4* rec;  // (pos, grho) per slot
  float px, py, pz, gi, d2, rd2, k6;
  float gpx, gpy, gpz;

  struct Rec {
    float4 a;
  };
  __device__ __forceinline__ Rec load(uint j) const { return Rec{rec[j]}; }
  __device__ __forceinline__ void eval(uint j, bool ok) { apply(load(j), ok); }
  __device__ __forceinline__ void apply(const Rec& r, bool ok) {
    const float4 a = r.a;
    const float dx = px - a.x, dy = py - a.y, dz = pz - a.z;
    const float dsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const float t = (rd2 - dsq) * d2;
    const float cf = ((gi + a.w) * (t * t)) * k6;
    const float nx = fmaf(cf, dx, gpx), ny = fmaf(cf, dy, gpy), nz = fmaf(cf, dz, gpz);
    if (ok) {
      gpx = nx; gpy = ny; gpz = nz;
    }
  }
};

extern "C" __global__ void __launch_bounds__(ADJ_THREADS) adjDensity(AdjArgs A) {
  __shared__ uint fq[(2 * AW_QD * 64 > AQ_SLOTS) ? 2 * AW_QD * 64 : AQ_SLOTS];
  const Slot S = slot_of(A);
  const FParams_t& P = *S.P;
  const int i = S.i;
  const int gc = i < S.p1 ? A.gcell[i] : FX_GRID_UNDEF;
  const bool act = gc != FX_GRID_UNDEF;
  const int base = gc - (P.gridRes.z + 1) * P.gridRes.x - 1;
  DensityAdj D;
  D.rec = A.drec;
  D.d2 = P.d2;
  D.rd2 = P.rd2;
  D.k6 = ((-6.0f * P.pmass) * P.poly6kern) * P.d2;
  D.px = D.py = D.pz = D.gi = 0.0f;
  D.gpx = D.gpy = D.gpz = 0.0f;
  if (act) {
    const float4 a = A.drec[i];
    D.px = a.x; D.py = a.y; D.pz = a.z;
    D.gi = a.w;
  }
  const bool ovf = pair_walk<AW_QD>(P, A, i, base, act, fq, D);
  if (ovf) {  // no recorded bits: the candidates in place
    for (int r = 0; r < P.gridSrch * P.gridSrch; r++) {
      int jb, je;
      row_of(P, base, r, A.off, A.cnt, jb, je);
      for (int j = jb; j < je; j++) {
        const float4 a = A.drec[j];
        const float dx = D.px - a.x, dy = D.py - a.y, dz = D.pz - a.z;
        const float dsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        if (dsq < D.rd2 && dsq > 0.0f) D.eval((uint)j, true);
      }
    }
  }
  if (i >= S.p1) return;
  const uint src = A.id[i];
  const fx_float3 x0 = A.gxs[i];
  A.gxp[src] = fx_float3{x0.x + D.gpx, x0.y + D.gpy, x0.z + D.gpz};
  A.gep[src] = A.ges[i];
}

// ----------------------------------------------------------------- reduce
// accum[3 s + k] (+)= sum of partial[3 b + k] over scene s's blocks
// [bfirst[s], bfirst[s+1]), k in [k0, k1), in a fixed order.  One block per scene.
extern "C" __global__ void __launch_bounds__(RED_THREADS)
    adjReduce(const double* partial, const int* bfirst, double* accum, int k0, int k1, int add) {
  __shared__ double lds[RED_THREADS / 64];
  const int s = blockIdx.x;
  const int b0 = bfirst[s], b1 = bfirst[s + 1];
  for (int k = k0; k < k1; k++) {
    double v = 0.0;
    for (int b = b0 + (int)threadIdx.x; b < b1; b += RED_THREADS) v += partial[3 * b + k];
    v = block_sum<RED_THREADS>(v, lds);
    if (threadIdx.x == 0) accum[3 * s + k] = add ? accum[3 * s + k] + v : v;
  }
}
