// oracle_capi.cpp -- C entry points of the CPU oracle (liboracle.so).
// TEST INFRASTRUCTURE: loaded only by tests/, the smoke check and the
// CPU-baseline leg of bench.py.
//
// Parity status of each function is pinned by tests/test_oracle.py:
//   orc_derive            grid + kernel constants   pinned: closed-form 16k/4M dam-break
//                                                     grid numbers, gridAdj formula
//   orc_run (insert)      cell, counts              pinned: hand-placed particles,
//                                                     brute-force counts
//   orc_run (scan)        offsets                   pinned: numpy cumsum
//   orc_run (sort)        permutation               pinned: numpy stable argsort
//   orc_run (pressure)    density                   pinned: FP64 brute force over all
//                                                     pairs, two-particle closed form,
//                                                     isolated particle = 1
//   orc_run (force)       force                     pinned: two-particle closed form,
//                                                     momentum sum, zero/uniform velocity
//   orc_run (advance)     integrator                pinned: free fall, each wall band,
//                                                     both limiters, GRID_UNDEF park
//   orc_pairs_*           neighbour counts          pinned: brute force
//   orc_sample            volume splat              pinned: single-particle closed form
//   orc_emit              emitter respawn           pinned: SplitMix64 known answer, distinct
//                                                     targets, box lattice, velocity
//   orc_trig              the wall's sine/cosine    pinned: against double-precision sin
//                                                     and cos within 2 ulps
//   orc_grad_*            adjoint of a rollout      pinned (FP64): central differences of
//                                                     orc_loss_f64 in every input kind,
//                                                     a two-particle closed form
#include <chrono>
#include <cmath>
#include <cstring>

#include "sph_adjoint.hpp"

namespace {

template <class R>
void load(oracle::State<R>& S, const FParams_t* P, int n, const R* pos, const R* vel, const R* veval,
          const R* force, const R* press, const uint32_t* clr, const uint32_t* id, const int32_t* gcell,
          const int32_t* gndx, const uint32_t* cnt, const uint32_t* off, const uint32_t* grid, R time) {
  S.resize(n, P->gridTotal);
  for (int i = 0; i < n; i++) {
    S.pos[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    S.vel[i] = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
    S.veval[i] = {veval[3 * i], veval[3 * i + 1], veval[3 * i + 2]};
    S.force[i] = {force[3 * i], force[3 * i + 1], force[3 * i + 2]};
    S.press[i] = press[i];
    S.clr[i] = clr[i];
    S.id[i] = id[i];
    S.gcell[i] = gcell[i];
    S.gndx[i] = gndx[i];
    S.grid[i] = grid[i];
  }
  memcpy(S.cnt.data(), cnt, sizeof(uint32_t) * (P->gridTotal + 1));
  memcpy(S.off.data(), off, sizeof(uint32_t) * (P->gridTotal + 1));
  S.time = time;
}

template <class R>
void store(const oracle::State<R>& S, R* pos, R* vel, R* veval, R* force, R* press, uint32_t* clr,
           uint32_t* id, int32_t* gcell, int32_t* gndx, uint32_t* cnt, uint32_t* off, uint32_t* grid,
           R* time) {
  for (int i = 0; i < S.n; i++) {
    pos[3 * i] = S.pos[i].x; pos[3 * i + 1] = S.pos[i].y; pos[3 * i + 2] = S.pos[i].z;
    vel[3 * i] = S.vel[i].x; vel[3 * i + 1] = S.vel[i].y; vel[3 * i + 2] = S.vel[i].z;
    veval[3 * i] = S.veval[i].x; veval[3 * i + 1] = S.veval[i].y; veval[3 * i + 2] = S.veval[i].z;
    force[3 * i] = S.force[i].x; force[3 * i + 1] = S.force[i].y; force[3 * i + 2] = S.force[i].z;
    press[i] = S.press[i];
    clr[i] = S.clr[i];
    id[i] = S.id[i];
    gcell[i] = S.gcell[i];
    gndx[i] = S.gndx[i];
    grid[i] = S.grid[i];
  }
  memcpy(cnt, S.cnt.data(), sizeof(uint32_t) * (S.G + 1));
  memcpy(off, S.off.data(), sizeof(uint32_t) * (S.G + 1));
  *time = S.time;
}

// stage bits: 1 insert, 2 scan, 4 sort, 8 pressure, 16 force, 32 advance
template <class R>
void run(const FParams_t* P, oracle::State<R>& S, int nsteps, int stage_mask, int threads) {
  for (int s = 0; s < nsteps; s++) {
    if (stage_mask & 1) oracle::insert(*P, S);
    if (stage_mask & 2) oracle::scan(S);
    if (stage_mask & 4) oracle::sort(*P, S, threads);
    if (stage_mask & 8) oracle::pressure(*P, S, threads);
    if (stage_mask & 16) oracle::force(*P, S, threads);
    if (stage_mask & 32) {
      oracle::advance(*P, S, threads);
      S.time = S.time + (R)P->dt;
    }
  }
}

// Independent restatement of Particles::RebuildAccelGrid
// (fluids5.0/src/particles.cpp:551-580,614-618) and UpdateParams (:659-673).
void derive(FParams_t* P) {
  P->grid_size = (float)(1.5 * (double)P->psmoothradius / (double)P->grid_density);
  const float g2 = (float)(2.0 * (double)(P->grid_size / P->sim_scale));
  P->gridMin = {P->bound_min.x - g2, P->bound_min.y - g2, P->bound_min.z - g2};
  P->gridMax = {P->bound_max.x + g2, P->bound_max.y + g2, P->bound_max.z + g2};
  const float sx = P->gridMax.x - P->gridMin.x, sy = P->gridMax.y - P->gridMin.y,
              sz = P->gridMax.z - P->gridMin.z;
  const float w = P->grid_size / P->sim_scale;
  P->gridRes = {(int)ceilf(sx / w), (int)ceilf(sy / w), (int)ceilf(sz / w)};
  P->gridSize = {(float)P->gridRes.x * w, (float)P->gridRes.y * w, (float)P->gridRes.z * w};
  P->gridDelta = {(float)P->gridRes.x / P->gridSize.x, (float)P->gridRes.y / P->gridSize.y,
                  (float)P->gridRes.z / P->gridSize.z};
  P->gridTotal = P->gridRes.x * P->gridRes.y * P->gridRes.z;
  int srch = (int)(floorf(2.0f * (P->psmoothradius / P->sim_scale) / w) + 1.0f);
  P->gridSrch = srch < 2 ? 2 : srch;
  P->gridAdjCnt = P->gridSrch * P->gridSrch * P->gridSrch;
  P->gridScanMax = {P->gridRes.x - P->gridSrch, P->gridRes.y - P->gridSrch, P->gridRes.z - P->gridSrch};
  int k = 0;
  if (P->gridSrch > 4) P->gridSrch = 0;  // gridAdj[64] cannot hold the stencil: mark invalid
  for (int y = 0; y < P->gridSrch; y++)
    for (int z = 0; z < P->gridSrch; z++)
      for (int x = 0; x < P->gridSrch; x++) P->gridAdj[k++] = (y * P->gridRes.z + z) * P->gridRes.x + x;
  P->gravity = {P->grav_dir.x * P->grav_amt, P->grav_dir.y * P->grav_amt, P->grav_dir.z * P->grav_amt};
  P->AL2 = P->AL * P->AL;
  P->VL2 = P->VL * P->VL;
  const float sr = P->psmoothradius;
  P->r2 = sr * sr;
  P->pdist = powf(P->pmass / P->prest_dens, 1 / 3.0f);
  P->poly6kern = 315.0f / (64.0f * 3.141592f * powf(sr, 9.0f));
  P->spikykern = -45.0f / (3.141592f * powf(sr, 6.0f));
  P->lapkern = 45.0f / (3.141592f * powf(sr, 6.0f));
  P->gausskern = 1.0f / powf(3.141592f * 2.0f * sr * sr, 3.0f / 2.0f);
  P->d2 = P->sim_scale * P->sim_scale;
  P->rd2 = P->r2 / P->d2;
  P->vterm = P->lapkern * P->pvisc;
}

}  // namespace

extern "C" {

void orc_derive(FParams_t* P) { derive(P); }

#define ORC_RUN(NAME, R)                                                                              \
  int NAME(const FParams_t* P, int n, R* pos, R* vel, R* veval, R* force, R* press, uint32_t* clr,  \
           uint32_t* id, int32_t* gcell, int32_t* gndx, uint32_t* cnt, uint32_t* off, uint32_t* grid, \
           R* time, int nsteps, int stage_mask, int threads) {                                        \
    oracle::State<R> S;                                                                               \
    load(S, P, n, pos, vel, veval, force, press, clr, id, gcell, gndx, cnt, off, grid, *time);        \
    run(P, S, nsteps, stage_mask, threads);                                                           \
    store(S, pos, vel, veval, force, press, clr, id, gcell, gndx, cnt, off, grid, time);              \
    return 0;                                                                                         \
  }
ORC_RUN(orc_run_f32, float)
ORC_RUN(orc_run_f64, double)

// Per-particle candidate count (members of the stencil cells) and in-radius
// pair count (0 < d2 < rd2) through the grid, on a cell-sorted state.
int orc_pairs_grid(const FParams_t* P, int n, const float* pos, const int32_t* gcell, const uint32_t* cnt,
                   const uint32_t* off, int32_t* cand, int32_t* pairs) {
  for (int i = 0; i < n; i++) {
    cand[i] = pairs[i] = 0;
    if (gcell[i] == FX_GRID_UNDEF) continue;
    const int base = gcell[i] - (P->gridRes.z + 1) * P->gridRes.x - 1;
    for (int r = 0; r < P->gridSrch * P->gridSrch; r++) {
      const int c0 = base + P->gridAdj[r * P->gridSrch], c1 = c0 + P->gridSrch - 1;
      for (uint32_t j = off[c0]; j < off[c1] + cnt[c1]; j++) {
        const float dx = pos[3 * i] - pos[3 * j], dy = pos[3 * i + 1] - pos[3 * j + 1],
                    dz = pos[3 * i + 2] - pos[3 * j + 2];
        const float dsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        cand[i]++;
        if (dsq < P->rd2 && dsq > 0.0f) pairs[i]++;
      }
    }
  }
  return 0;
}

// The same in-radius pair count and an FP64 density by brute force over all
// pairs (no grid), for small n.
int orc_pairs_brute(const FParams_t* P, int n, const float* pos, int32_t* pairs, double* density) {
  const double rd2 = P->rd2, d2 = P->d2;
  for (int i = 0; i < n; i++) {
    int np = 0;
    double sum = 0;
    for (int j = 0; j < n; j++) {
      const float dx = pos[3 * i] - pos[3 * j], dy = pos[3 * i + 1] - pos[3 * j + 1],
                  dz = pos[3 * i + 2] - pos[3 * j + 2];
      const float dsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
      if (dsq < P->rd2 && dsq > 0.0f) {
        np++;
        const double ddx = (double)pos[3 * i] - pos[3 * j], ddy = (double)pos[3 * i + 1] - pos[3 * j + 1],
                     ddz = (double)pos[3 * i + 2] - pos[3 * j + 2];
        const double t = (rd2 - (ddx * ddx + ddy * ddy + ddz * ddz)) * d2;
        sum += t * t * t;
      }
    }
    pairs[i] = np;
    double rho = sum * (double)P->pmass * (double)P->poly6kern;
    density[i] = rho == 0.0 ? 1.0 : rho;
  }
  return 0;
}

// Wall-clock seconds of nsteps full FP32 steps on threads host threads
// (the CPU baseline); the state is advanced in place.
double orc_time_f32(const FParams_t* P, int n, float* pos, float* vel, float* veval, float* force,
                    float* press, uint32_t* clr, uint32_t* id, int32_t* gcell, int32_t* gndx,
                    uint32_t* cnt, uint32_t* off, uint32_t* grid, float* time, int nsteps, int threads) {
  oracle::State<float> S;
  load(S, P, n, pos, vel, veval, force, press, clr, id, gcell, gndx, cnt, off, grid, *time);
  const auto t0 = std::chrono::steady_clock::now();
  for (int s = 0; s < nsteps; s++) oracle::step(*P, S, threads);
  const auto t1 = std::chrono::steady_clock::now();
  store(S, pos, vel, veval, force, press, clr, id, gcell, gndx, cnt, off, grid, time);
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"

// ---------------------------------------------------------------- adjoint
namespace {

template <class R>
void load_initial(oracle::State<R>& S, const FParams_t* P, int n, const R* pos, const R* vel, const R* veval,
                  const uint32_t* clr, const uint32_t* id, R time, int use_coef, double visc, double k) {
  S.resize(n, P->gridTotal);
  for (int i = 0; i < n; i++) {
    S.pos[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    S.vel[i] = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
    S.veval[i] = {veval[3 * i], veval[3 * i + 1], veval[3 * i + 2]};
    S.clr[i] = clr[i];
    S.id[i] = id[i];
  }
  S.time = time;
  if (use_coef) {
    S.coef_set = true;
    S.vterm = (R)P->lapkern * (R)visc;
    S.stiff = (R)k;
  }
}

}  // namespace

extern "C" {

// Loss of an nsteps rollout and its adjoint: gradients of the loss with
// respect to the initial pos, vel and veval of every particle (by id, 3n each),
// viscosity (through vterm = lapkern visc) and internal stiffness.  With
// use_coef the force uses visc and k in the state's precision, otherwise
// the FParams_t values.
#define ORC_GRAD(NAME, R)                                                                                 \
  int NAME(const FParams_t* P, int n, const R* pos, const R* vel, const R* veval, const uint32_t* clr,  \
           const uint32_t* id, R time, int nsteps, int use_coef, double visc, double k, const R* target,   \
           int threads, double* loss, R* gpos, R* gvel, R* gveval, double* gvisc, double* gk) {          \
    oracle::State<R> S;                                                                                   \
    load_initial(S, P, n, pos, vel, veval, clr, id, time, use_coef, visc, k);                             \
    oracle::Adjoint<R> G;                                                                                 \
    *loss = oracle::loss_and_grad(*P, S, nsteps, target, G, threads);                                     \
    for (int i = 0; i < n; i++) {                                                                         \
      gpos[3 * i] = G.gx[i].x; gpos[3 * i + 1] = G.gx[i].y; gpos[3 * i + 2] = G.gx[i].z;                  \
      gvel[3 * i] = G.gv[i].x; gvel[3 * i + 1] = G.gv[i].y; gvel[3 * i + 2] = G.gv[i].z;                  \
      gveval[3 * i] = G.ge[i].x; gveval[3 * i + 1] = G.ge[i].y; gveval[3 * i + 2] = G.ge[i].z;            \
    }                                                                                                     \
    *gvisc = G.gvisc;                                                                                      \
    *gk = G.gk;                                                                                           \
    return 0;                                                                                             \
  }
ORC_GRAD(orc_grad_f32, float)
ORC_GRAD(orc_grad_f64, double)

// FP32 rollout, FP64 adjoint over its (promoted) tape; outputs in double.
int orc_grad_mixed(const FParams_t* P, int n, const float* pos, const float* vel, const float* veval,
                   const uint32_t* clr, const uint32_t* id, float time, int nsteps, const float* target,
                   int threads, double* loss, double* gpos, double* gvel, double* gveval, double* gvisc,
                   double* gk) {
  oracle::State<float> S;
  load_initial(S, P, n, pos, vel, veval, clr, id, time, 0, 0.0, 0.0);
  oracle::Adjoint<double> G;
  *loss = oracle::loss_and_grad_mixed(*P, S, nsteps, target, G, threads);
  for (int i = 0; i < n; i++) {
    gpos[3 * i] = G.gx[i].x; gpos[3 * i + 1] = G.gx[i].y; gpos[3 * i + 2] = G.gx[i].z;
    gvel[3 * i] = G.gv[i].x; gvel[3 * i + 1] = G.gv[i].y; gvel[3 * i + 2] = G.gv[i].z;
    gveval[3 * i] = G.ge[i].x; gveval[3 * i + 1] = G.ge[i].y; gveval[3 * i + 2] = G.ge[i].z;
  }
  *gvisc = G.gvisc;
  *gk = G.gk;
  return 0;
}

// SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014): one step from state x.
uint64_t orc_splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// emitParticles (fluids4.0/fluid_kernels.cu:231-257) with the draw of the
// GPU kernel: thread i < min(emit, n) takes v = high word of
// splitmix(h ^ splitmix(i)), h = splitmix(bits of frame), and respawns
// particle (2654435761 i + h mod n) mod n at bmin + (byte/256)(bmax - bmin)
// per axis with velocity (5,-2,0), zero veval and force.
int orc_emit(int n, float frame, int emit, float* pos, float* vel, float* veval, float* force) {
  uint32_t fb;
  std::memcpy(&fb, &frame, 4);
  const uint64_t h = orc_splitmix64((uint64_t)fb);
  const int m = emit < n ? emit : n;
  const float bmin[3] = {-170.0f, 10.0f, -20.0f}, bmax[3] = {-190.0f, 60.0f, 20.0f};
  for (int i = 0; i < m; i++) {
    const uint32_t v = (uint32_t)(orc_splitmix64(h ^ orc_splitmix64((uint64_t)i)) >> 32);
    const uint64_t j = (2654435761ull * (uint64_t)i + h % (uint64_t)n) % (uint64_t)n;
    for (int a = 0; a < 3; a++) {
      const float u = (float)((v >> (8 * a)) & 0xFFu) / 256.0f;
      const float t = u * (bmax[a] - bmin[a]);
      pos[3 * j + a] = bmin[a] + t;
      veval[3 * j + a] = 0.0f;
      force[3 * j + a] = 0.0f;
    }
    vel[3 * j] = 5.0f;
    vel[3 * j + 1] = -2.0f;
    vel[3 * j + 2] = 0.0f;
  }
  return 0;
}

// The moving wall's sine and cosine (fx_trig.hpp, the kernels' own) of n
// floats, for the accuracy check.
int orc_trig(int n, const float* x, float* s, float* c) {
  for (int i = 0; i < n; i++) {
    s[i] = fx_sinf(x[i]);
    c[i] = fx_cosf(x[i]);
  }
  return 0;
}

// sampleParticles (fluids4.0/fluid_kernels.cu:259-300) on a cell-sorted
// state: brick[(z*ry + y)*rx + x], voxels outside the scanned cell range 0.
int orc_sample(const FParams_t* P, int n, const float* pos, const uint32_t* cnt, const uint32_t* off,
               const int32_t* res, const float* bmin, const float* bmax, float scalar, float* brick) {
  const float h2 = 2.0f * P->r2 / 8.0f;
  const int nadj = (P->gridRes.z + 1) * P->gridRes.x + 1;
  for (int z = 0; z < res[2]; z++)
    for (int y = 0; y < res[1]; y++)
      for (int x = 0; x < res[0]; x++) {
        const float px = bmin[0] + ((float)x / (float)res[0]) * (bmax[0] - bmin[0]);
        const float py = bmin[1] + ((float)y / (float)res[1]) * (bmax[1] - bmin[1]);
        const float pz = bmin[2] + ((float)z / (float)res[2]) * (bmax[2] - bmin[2]);
        const int gx = (int)((px - P->gridMin.x) * P->gridDelta.x);
        const int gy = (int)((py - P->gridMin.y) * P->gridDelta.y);
        const int gz = (int)((pz - P->gridMin.z) * P->gridDelta.z);
        float v = 0.0f;
        const bool in = !(gx < 1 || gx > P->gridRes.x - P->gridSrch || gy < 1 || gy > P->gridRes.y - P->gridSrch ||
                          gz < 1 || gz > P->gridRes.z - P->gridSrch);
        if (in) {
          const int gs = (gy * P->gridRes.z + gz) * P->gridRes.x + gx - nadj;
          for (int c = 0; c < P->gridAdjCnt; c++) {
            const int cell = gs + P->gridAdj[c];
            for (int j = (int)off[cell]; j < (int)(off[cell] + cnt[cell]) && j < n; j++) {
              const float dx = px - pos[3 * j], dy = py - pos[3 * j + 1], dz = pz - pos[3 * j + 2];
              const float dsq = dx * dx + dy * dy + dz * dz;
              if (dsq < P->rd2 && dsq > 0.0f) {
                const float r = std::sqrt(dsq * P->d2);
                v += P->gausskern * std::exp(-(r * r) / h2);
              }
            }
          }
          v = v * scalar;
        }
        brick[((size_t)z * res[1] + y) * res[0] + x] = v;
      }
  return 0;
}

#define ORC_LOSS(NAME, R)                                                                                 \
  double NAME(const FParams_t* P, int n, const R* pos, const R* vel, const R* veval, const uint32_t* clr, \
              const uint32_t* id, R time, int nsteps, int use_coef, double visc, double k, const R* target, \
              int threads) {                                                                              \
    oracle::State<R> S;                                                                                   \
    load_initial(S, P, n, pos, vel, veval, clr, id, time, use_coef, visc, k);                             \
    oracle::rollout(*P, S, nsteps, (std::vector<oracle::TapeStep<R>>*)nullptr, threads);                  \
    oracle::Adjoint<R> G;                                                                                 \
    return oracle::loss_seed(S, target, G);                                                               \
  }
ORC_LOSS(orc_loss_f32, float)
ORC_LOSS(orc_loss_f64, double)

}  // extern "C"
