// sph_oracle.hpp -- CPU oracle of the SPH step.  TEST INFRASTRUCTURE: used
// only by tests/, the smoke check and the CPU-baseline leg of bench.py, as the
// checker; never by the product path.
//
// A restatement of the reference GPU step (the normative semantics,
// fluids5.0/src/particles.cu) with the loop structure of the reference's only
// complete CPU pipeline (fluids3.9/fluid_system.cpp:335-346,462-639):
// sequential insert (so the in-cell order is the particle order), CPU
// exclusive scan, counting sort, density, force, integrator.
//
// Templated on Real: Real=float evaluates the same IEEE-754 single-precision
// expression trees as the gfx950 kernels (explicit fused multiply-adds at the
// same places, no other contraction: build with -ffp-contract=off), Real=double
// is the FP64 build of the same formulas used to state the FP32 error.
//
// Project choices shared with the kernels (documented divergences from the
// reference):
//   - out-of-range particles are counted in slot G (= gridTotal) and sorted
//     into the tail, in particle order;
//   - press and force are not carried through the sort (the next two stages
//     overwrite them for every in-range particle).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/fluidsx_params.h"
#include "../fluids3_amd/csrc/fx_trig.hpp"

namespace oracle {

// Sine and cosine of the moving wall: the float step takes the same written-out
// sequence of IEEE operations as the kernels (fluids3_amd/csrc/fx_trig.hpp), so that both
// produce the same bits; the double step takes the C library's.
inline float orc_sin(float x) { return fx_sinf(x); }
inline double orc_sin(double x) { return std::sin(x); }
inline float orc_cos(float x) { return fx_cosf(x); }
inline double orc_cos(double x) { return std::cos(x); }

template <class R>
struct V3 {
  R x, y, z;
};

template <class R>
struct State {
  int n = 0;
  int G = 0;  // gridTotal
  std::vector<V3<R>> pos, vel, veval, force;
  std::vector<R> press;
  std::vector<uint32_t> clr, id;
  std::vector<int32_t> gcell, gndx;
  std::vector<uint32_t> cnt, off;  // G+1 entries
  std::vector<uint32_t> grid;      // member list of each cell (index sort)
  R time = 0;
  // force coefficients in the state's precision when set (the FP64 rollouts
  // of the adjoint checks perturb viscosity and stiffness below float
  // resolution); otherwise the float FParams_t values
  bool coef_set = false;
  R vterm = 0, stiff = 0;

  void resize(int n_, int G_) {
    n = n_;
    G = G_;
    pos.assign(n, V3<R>{0, 0, 0});
    vel = veval = force = pos;
    press.assign(n, 0);
    clr.assign(n, 0);
    id.assign(n, 0);
    gcell.assign(n, 0);
    gndx.assign(n, 0);
    cnt.assign(G + 1, 0);
    off.assign(G + 1, 0);
    grid.assign(n, 0);
  }
};

// --------------------------------------------------------------- insert
// particles.cu:24-54 (cell formula, range test), fluids3.9
// fluid_system.cpp:462-495 (sequential counting).  The range test is taken on
// the float cell coordinate: trunc(g) >= 1 <=> g >= 1, trunc(g) <= M <=> g < M+1.
template <class R>
void insert(const FParams_t& P, State<R>& S) {
  std::fill(S.cnt.begin(), S.cnt.end(), 0u);
  for (int i = 0; i < S.n; i++) {
    // the cell is always computed in float, as on the device
    const float px = (float)S.pos[i].x, py = (float)S.pos[i].y, pz = (float)S.pos[i].z;
    const float gx = (px - P.gridMin.x) * P.gridDelta.x;
    const float gy = (py - P.gridMin.y) * P.gridDelta.y;
    const float gz = (pz - P.gridMin.z) * P.gridDelta.z;
    const bool in = gx >= 1.0f && gx < (float)(P.gridScanMax.x + 1) && gy >= 1.0f &&
                    gy < (float)(P.gridScanMax.y + 1) && gz >= 1.0f && gz < (float)(P.gridScanMax.z + 1);
    int slot;
    if (in) {
      slot = ((int)gy * P.gridRes.z + (int)gz) * P.gridRes.x + (int)gx;
      S.gcell[i] = slot;
    } else {
      slot = S.G;
      S.gcell[i] = FX_GRID_UNDEF;
    }
    S.gndx[i] = (int32_t)S.cnt[slot]++;
  }
}

// ----------------------------------------------------------------- scan
// off[c] = sum_{c' < c} cnt[c'] (particles.cpp:818-827, the CPU branch)
template <class R>
void scan(State<R>& S) {
  uint32_t sum = 0;
  for (int c = 0; c <= S.G; c++) {
    S.off[c] = sum;
    sum += S.cnt[c];
  }
}

// ----------------------------------------------------------------- sort
// fluids3.9 fluid_system.cpp:496-517 / particles.cu:64-95: scatter to
// off[cell] + rank; grid[] receives the pre-sort index (index sort,
// fluids3.2/fluids/fluid_system_kern.cu:139-150).  The reference's in-cell
// order is the insert's atomic order (particles.cu:41), which no run
// reproduces; this step fixes it: the members of a cell are ordered by their
// x slab (the cell split into FX_SLABS = 16 equal slabs along x, slab_of below),
// ties by pre-sort index, and gndx becomes that rank.  Cells of more than
// FX_RANK_CAP members and the out-of-range tail keep particle order.
#ifndef FX_RANK_CAP
#define FX_RANK_CAP 65536
#endif

// x slab of an in-range position inside its cell, from the insert's float
// cell coordinate gx: floor(frac(gx) * FX_SLABS) (gx - (int)gx is exact for
// gx >= 1, and the product by a power of two is exact).
inline int slab_of(const FParams_t& P, float px) {
  const float gx = (px - P.gridMin.x) * P.gridDelta.x;
  const float f = gx - (float)(int)gx;
  const int s = (int)(f * (float)FX_SLABS);
  return s < FX_SLABS - 1 ? s : FX_SLABS - 1;
}

template <class R>
void sort(const FParams_t& P, State<R>& S, int threads = 1) {
  const State<R> T = S;  // the reference's CopyAllToTemp
  // member lists in particle order (the sequential insert's gndx)
  std::vector<uint32_t> mem(S.n);
  for (int i = 0; i < S.n; i++) {
    const int c = T.gcell[i];
    const int cell = (c == FX_GRID_UNDEF) ? S.G : c;
    mem[S.off[cell] + (uint32_t)T.gndx[i]] = (uint32_t)i;
  }
#pragma omp parallel for schedule(dynamic, 1024) num_threads(threads) if (threads > 1)
  for (int c = 0; c < S.G; c++) {
    const uint32_t n = S.cnt[c];
    if (n < 2 || n > FX_RANK_CAP) continue;
    uint32_t* m = mem.data() + S.off[c];
    std::sort(m, m + n, [&](uint32_t a, uint32_t b) {
      const int sa = slab_of(P, (float)T.pos[a].x), sb = slab_of(P, (float)T.pos[b].x);
      return sa != sb ? sa < sb : a < b;
    });
  }
  for (int c = 0; c <= S.G; c++) {
    for (uint32_t r = 0; r < S.cnt[c]; r++) {
      const uint32_t s = S.off[c] + r, i = mem[s];
      S.pos[s] = T.pos[i];
      S.vel[s] = T.vel[i];
      S.veval[s] = T.veval[i];
      S.clr[s] = T.clr[i];
      S.id[s] = T.id[i];
      S.gcell[s] = T.gcell[i];
      S.gndx[s] = (int32_t)r;
      S.grid[s] = i;
    }
  }
}

// stencil row r of cell gc: sorted index range [jb, je)
template <class R>
inline void row_range(const FParams_t& P, const State<R>& S, int base, int r, int& jb, int& je) {
  const int srch = P.gridSrch;
  const int c0 = base + P.gridAdj[r * srch];
  const int c1 = c0 + srch - 1;
  jb = (int)S.off[c0];
  je = (int)(S.off[c1] + S.cnt[c1]);
}

// ------------------------------------------------------------- density
// particles.cu:120-157 (same formula in fluids3.9 fluid_system.cpp:521-570)
template <class R>
void pressure_one(const FParams_t& P, State<R>& S, int i) {
  const int gc = S.gcell[i];
  if (gc == FX_GRID_UNDEF) return;
  const int base = gc - (P.gridRes.z + 1) * P.gridRes.x - 1;
  const R rd2 = P.rd2, d2 = P.d2;
  const V3<R> p = S.pos[i];
  R sum = 0;
  for (int r = 0; r < P.gridSrch * P.gridSrch; r++) {
    int jb, je;
    row_range(P, S, base, r, jb, je);
    for (int j = jb; j < je; j++) {
      const V3<R> q = S.pos[j];
      const R dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
      const R dsq = std::fma(dz, dz, std::fma(dy, dy, dx * dx));
      if (dsq < rd2 && dsq > R(0)) {
        const R t = (rd2 - dsq) * d2;
        sum = std::fma(t * t, t, sum);
      }
    }
  }
  sum = sum * (R)P.pmass * (R)P.poly6kern;
  if (sum == R(0)) sum = R(1);
  S.press[i] = sum;
}

// --------------------------------------------------------------- force
// particles.cu:188-225 (fluids3.9 fluid_system.cpp:573-631):
//   F_i += (pterm dist + vterm (veval_j - veval_i)) c / (rho_i rho_j),
//   pterm = ss (-0.5) c spiky (p_i + p_j) / r,  c = h - r,  p = (rho - rho0) k.
// Evaluated with one reciprocal: inv = 1 / (r rho_i rho_j),
//   A = ((ss -0.5) spiky) c^2 (p_i + p_j) inv,  B = vterm c r inv,
//   F += A dist + B (veval_j - veval_i).
template <class R>
void force_one(const FParams_t& P, State<R>& S, int i) {
  const int gc = S.gcell[i];
  if (gc == FX_GRID_UNDEF) return;
  const int base = gc - (P.gridRes.z + 1) * P.gridRes.x - 1;
  const R rd2 = P.rd2, d2 = P.d2, h = P.psmoothradius;
  const R rest = P.prest_dens;
  const R stiff = S.coef_set ? S.stiff : (R)P.pintstiff;
  const R kp = ((R)P.sim_scale * R(-0.5)) * (R)P.spikykern;
  const R vterm = S.coef_set ? S.vterm : (R)P.vterm;
  const V3<R> p = S.pos[i];
  const V3<R> vi = S.veval[i];
  const R di = S.press[i];
  const R pi = (di - rest) * stiff;
  R fx = 0, fy = 0, fz = 0;
  for (int r = 0; r < P.gridSrch * P.gridSrch; r++) {
    int jb, je;
    row_range(P, S, base, r, jb, je);
    for (int j = jb; j < je; j++) {
      const V3<R> q = S.pos[j];
      const R dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
      const R dsq = std::fma(dz, dz, std::fma(dy, dy, dx * dx));
      if (dsq < rd2 && dsq > R(0)) {
        const R rr = std::sqrt(dsq * d2);
        const R c = h - rr;
        const R dj = S.press[j];
        const R pj = (dj - rest) * stiff;
        const R inv = R(1) / (rr * (di * dj));
        const R A = (kp * (c * c)) * (pi + pj) * inv;
        const R B = (vterm * c) * rr * inv;
        const V3<R> vj = S.veval[j];
        fx = std::fma(A, dx, std::fma(B, vj.x - vi.x, fx));
        fy = std::fma(A, dy, std::fma(B, vj.y - vi.y, fy));
        fz = std::fma(A, dz, std::fma(B, vj.z - vi.z, fz));
      }
    }
  }
  S.force[i] = V3<R>{fx, fy, fz};
}

// ------------------------------------------------------------- advance
// particles.cu:237-322 (walls on bound_min/max, GPU variant; not the
// gridMin/Max walls of the reference CPU Advance).
template <class R>
void advance_one(const FParams_t& P, State<R>& S, int i, R time, R dt, R ss) {
  if (S.gcell[i] == FX_GRID_UNDEF) {
    S.pos[i] = V3<R>{R(-1000), R(-1000), R(-1000)};
    S.vel[i] = V3<R>{0, 0, 0};
    return;
  }
  const V3<R> pos = S.pos[i];
  V3<R> vel = S.vel[i];
  const V3<R> ve = S.veval[i];
  const V3<R> f = S.force[i];
  const R m = P.pmass;
  R ax = f.x * m, ay = f.y * m, az = f.z * m;
  const R rad = P.pradius, stiff = P.bound_stiff, damp = P.bound_damp;
  const R slope = P.bound_slope, fric = P.bound_friction;
  const R bminx = P.bound_min.x, bminy = P.bound_min.y, bminz = P.bound_min.z;
  const R bmaxx = P.bound_max.x, bmaxy = P.bound_max.y, bmaxz = P.bound_max.z;
  const R eps = (R)FX_EPSILON;
  R diff, adj, nx, ny, nz;

  diff = rad - (pos.y - (bminy + (pos.x - bminx) * slope)) * ss;
  if (diff > eps) {
    nx = -slope;
    ny = R(1) - slope;
    nz = 0;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    nx *= adj; ny *= adj; nz *= adj;
    ax += nx - ve.x * fric;
    ay += ny - ve.y * fric;
    az += nz - ve.z * fric;
  }
  diff = rad - (bmaxy - pos.y) * ss;
  if (diff > eps) {
    nx = 0; ny = -1; nz = 0;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    nx *= adj; ny *= adj; nz *= adj;
    ax += nx - ve.x * fric;
    ay += ny - ve.y * fric;
    az += nz - ve.z * fric;
  }
  const R sarg = time * (R)P.bound_wall_freq + pos.z / R(400);
  const R wall = (R)(((double)orc_sin(sarg) * 0.5 + 0.5) * (double)P.bound_wall_force);
  diff = rad - (pos.x - (bminx + wall)) * ss;
  if (diff > eps) {
    nx = 1; ny = 0; nz = 0;
    adj = wall * stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    ax += nx * adj; ay += ny * adj; az += nz * adj;
  }
  diff = rad - (bmaxx - pos.x) * ss;
  if (diff > eps) {
    nx = -1; ny = 0; nz = 0;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    ax += nx * adj; ay += ny * adj; az += nz * adj;
  }
  diff = rad - (pos.z - bminz) * ss;
  if (diff > eps) {
    nx = 0; ny = 0; nz = 1;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    ax += nx * adj; ay += ny * adj; az += nz * adj;
  }
  diff = rad - (bmaxz - pos.z) * ss;
  if (diff > eps) {
    nx = 0; ny = 0; nz = -1;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    ax += nx * adj; ay += ny * adj; az += nz * adj;
  }
  ax += (R)P.gravity.x;
  ay += (R)P.gravity.y;
  az += (R)P.gravity.z;
  R speed = ax * ax + ay * ay + az * az;
  if (speed > (R)P.AL2) {
    const R s = (R)P.AL / std::sqrt(speed);
    ax *= s; ay *= s; az *= s;
  }
  speed = vel.x * vel.x + vel.y * vel.y + vel.z * vel.z;
  if (speed > (R)P.VL2) {
    const R s = (R)P.VL / std::sqrt(speed);
    vel.x *= s; vel.y *= s; vel.z *= s;
  }
  const R vnx = ax * dt + vel.x, vny = ay * dt + vel.y, vnz = az * dt + vel.z;
  S.veval[i] = V3<R>{(vel.x + vnx) * R(0.5), (vel.y + vny) * R(0.5), (vel.z + vnz) * R(0.5)};
  S.vel[i] = V3<R>{vnx, vny, vnz};
  const R dts = dt / ss;
  S.pos[i] = V3<R>{pos.x + vnx * dts, pos.y + vny * dts, pos.z + vnz * dts};
}

template <class R>
void pressure(const FParams_t& P, State<R>& S, int threads) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads) if (threads > 1)
  for (int i = 0; i < S.n; i++) pressure_one(P, S, i);
}

template <class R>
void force(const FParams_t& P, State<R>& S, int threads) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads) if (threads > 1)
  for (int i = 0; i < S.n; i++) force_one(P, S, i);
}

template <class R>
void advance(const FParams_t& P, State<R>& S, int threads) {
  const R time = S.time, dt = P.dt, ss = P.sim_scale;
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
  for (int i = 0; i < S.n; i++) advance_one(P, S, i, time, dt, ss);
}

// One Particles::Run (particles.cpp:382-394); time advances in the state's
// precision like the host's m_Params.time.
template <class R>
void step(const FParams_t& P, State<R>& S, int threads) {
  insert(P, S);
  scan(S);
  sort(P, S, threads);
  pressure(P, S, threads);
  force(P, S, threads);
  advance(P, S, threads);
  S.time = S.time + (R)P.dt;
}

}  // namespace oracle
