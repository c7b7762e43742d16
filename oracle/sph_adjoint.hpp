// sph_adjoint.hpp -- CPU oracle of the reverse-mode (adjoint) SPH rollout.
// TEST INFRASTRUCTURE: used only by tests/ and the checks, never by the
// product path.
//
// The forward is the oracle step (sph_oracle.hpp).  A rollout records, for
// every step, the cell-sorted state right after the force stage (positions,
// velocities, evaluation velocities, forces, densities, ids, cells, count and
// offset tables): the neighbour sets, cells, branch masks and the sort
// permutation of the forward are taken as piecewise-constant data.
//
// Adjoint variables live in the particle-id frame (index = original
// particle), so the sort's permutation is undone by indexing with the
// recorded ids.  One backward step, for the step's sorted slots:
//   1. advance:  (gx', gv', ge') of the step's outputs -> gF (force), partial
//      gx/ge of the step's sorted inputs, and gv (final, velocity input);
//   2. force:    gather over the in-radius neighbours j of i, with
//      dl = gF_i - gF_j, u = x_i - x_j, dve = ve_j - ve_i:
//        gx_i   += A dl + (d2/r) u (A_r (u.dl) + B_r (dve.dl))
//        ge_i   -= B dl
//        grho_i += (kp c^2 inv) k (u.dl) - (A (u.dl) + B (dve.dl)) / rho_i
//        dL/dk    += (kp c^2 inv) ((rho_i - rho0) + (rho_j - rho0)) (u.gF_i)
//        dL/dvisc += lap c / (rho_i rho_j) (dve.gF_i)
//      with A = kp c^2 (p_i + p_j) inv, B = vterm c / (rho_i rho_j),
//      inv = 1 / (r rho_i rho_j), A_r = dA/dr, B_r = dB/dr;
//   3. density:  gx_i += sum_j (grho_i + grho_j) (-6 m poly6 d2) t^2 u,
//      t = (rd2 - |u|^2) d2.
// Each pair term is written once per ordered pair from i's side (a gather,
// no scatter), using the symmetry of A and B in i and j.
//
// Templated on Real like the forward: Real=float evaluates the same
// expression trees as the gfx950 adjoint kernels (bit-identical per-particle
// results), Real=double is the FP64 adjoint of the FP64 forward.
#pragma once

#include "sph_oracle.hpp"

namespace oracle {

template <class R>
struct TapeStep {
  std::vector<V3<R>> pos, vel, veval, force;
  std::vector<R> press;
  std::vector<uint32_t> id;
  std::vector<int32_t> gcell;
  std::vector<uint32_t> cnt, off;
  R time = 0;
};

template <class R>
struct Adjoint {
  std::vector<V3<R>> gx, gv, ge;  // id frame
  double gvisc = 0, gk = 0;
};

template <class R>
void record(const State<R>& S, TapeStep<R>& T) {
  T.pos = S.pos;
  T.vel = S.vel;
  T.veval = S.veval;
  T.force = S.force;
  T.press = S.press;
  T.id = S.id;
  T.gcell = S.gcell;
  T.cnt = S.cnt;
  T.off = S.off;
  T.time = S.time;
}

// nsteps oracle steps, recording each step after its force stage
template <class R>
void rollout(const FParams_t& P, State<R>& S, int nsteps, std::vector<TapeStep<R>>* tape, int threads) {
  if (tape) tape->assign(nsteps, TapeStep<R>());
  for (int t = 0; t < nsteps; t++) {
    insert(P, S);
    scan(S);
    sort(P, S, threads);
    pressure(P, S, threads);
    force(P, S, threads);
    if (tape) record(S, (*tape)[t]);
    advance(P, S, threads);
    S.time = S.time + (R)P.dt;
  }
}

// L = (1/n) sum |x - x*|^2 over all particles (FP64 sum in slot order);
// seeds gx[id] = (2/n) (x - x*), gv = ge = 0.  target: 3n values by id.
template <class R>
double loss_seed(const State<R>& S, const R* target, Adjoint<R>& G) {
  const int n = S.n;
  G.gx.assign(n, V3<R>{0, 0, 0});
  G.gv = G.ge = G.gx;
  G.gvisc = G.gk = 0;
  const R scale = R(2) / (R)n;
  double L = 0;
  for (int s = 0; s < n; s++) {
    const uint32_t id = S.id[s];
    const R dx = S.pos[s].x - target[3 * id], dy = S.pos[s].y - target[3 * id + 1],
            dz = S.pos[s].z - target[3 * id + 2];
    L += (double)dx * (double)dx + (double)dy * (double)dy + (double)dz * (double)dz;
    G.gx[id] = V3<R>{dx * scale, dy * scale, dz * scale};
  }
  return L / (double)n;
}

// -------------------------------------------------------- advance (reverse)
// Reverse of advance_one (sph_oracle.hpp) for sorted slot s of step T.
// In: output adjoints gxo, gvo, geo.  Out: gF, partial gx/ge of the inputs,
// gvel (complete).
template <class R>
void advance_bwd_one(const FParams_t& P, const TapeStep<R>& T, int s, R dt, R ss, V3<R> gxo, V3<R> gvo,
                     V3<R> geo, V3<R>& gF, V3<R>& gx, V3<R>& ge, V3<R>& gvel) {
  if (T.gcell[s] == FX_GRID_UNDEF) {  // parked: pos and vel become constants, veval passes through
    gF = gx = gvel = V3<R>{0, 0, 0};
    ge = geo;
    return;
  }
  const V3<R> pos = T.pos[s], vel = T.vel[s], ve = T.veval[s], f = T.force[s];
  const R m = P.pmass;
  R ax = f.x * m, ay = f.y * m, az = f.z * m;
  const R rad = P.pradius, stiff = P.bound_stiff, damp = P.bound_damp;
  const R slope = P.bound_slope, fric = P.bound_friction;
  const R bminx = P.bound_min.x, bminy = P.bound_min.y, bminz = P.bound_min.z;
  const R bmaxx = P.bound_max.x, bmaxy = P.bound_max.y, bmaxz = P.bound_max.z;
  const R eps = (R)FX_EPSILON;
  R diff, adj, nx, ny, nz;
  // forward replay: pre-limit acceleration and the active walls
  bool w1, w2, w3, w4, w5, w6;
  diff = rad - (pos.y - (bminy + (pos.x - bminx) * slope)) * ss;
  w1 = diff > eps;
  if (w1) {
    nx = -slope; ny = R(1) - slope; nz = 0;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    nx *= adj; ny *= adj; nz *= adj;
    ax += nx - ve.x * fric; ay += ny - ve.y * fric; az += nz - ve.z * fric;
  }
  diff = rad - (bmaxy - pos.y) * ss;
  w2 = diff > eps;
  if (w2) {
    nx = 0; ny = -1; nz = 0;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    nx *= adj; ny *= adj; nz *= adj;
    ax += nx - ve.x * fric; ay += ny - ve.y * fric; az += nz - ve.z * fric;
  }
  const R sarg = T.time * (R)P.bound_wall_freq + pos.z / R(400);
  const R wall = (R)(((double)orc_sin(sarg) * 0.5 + 0.5) * (double)P.bound_wall_force);
  diff = rad - (pos.x - (bminx + wall)) * ss;
  const R diff3 = diff;
  w3 = diff > eps;
  if (w3) {
    nx = 1; ny = 0; nz = 0;
    adj = wall * stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    ax += nx * adj; ay += ny * adj; az += nz * adj;
  }
  diff = rad - (bmaxx - pos.x) * ss;
  w4 = diff > eps;
  if (w4) {
    nx = -1; ny = 0; nz = 0;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    ax += nx * adj; ay += ny * adj; az += nz * adj;
  }
  diff = rad - (pos.z - bminz) * ss;
  w5 = diff > eps;
  if (w5) {
    nx = 0; ny = 0; nz = 1;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    ax += nx * adj; ay += ny * adj; az += nz * adj;
  }
  diff = rad - (bmaxz - pos.z) * ss;
  w6 = diff > eps;
  if (w6) {
    nx = 0; ny = 0; nz = -1;
    adj = stiff * diff - damp * (nx * ve.x + ny * ve.y + nz * ve.z);
    ax += nx * adj; ay += ny * adj; az += nz * adj;
  }
  ax += (R)P.gravity.x;
  ay += (R)P.gravity.y;
  az += (R)P.gravity.z;
  const R sa2 = ax * ax + ay * ay + az * az;
  const R sv2 = vel.x * vel.x + vel.y * vel.y + vel.z * vel.z;

  // reverse: x' = x + vn dt/ss, v' = vn, ve' = (v_l + vn) / 2, vn = a_l dt + v_l
  const R dts = dt / ss;
  const R h = R(0.5);
  const R gnx = gxo.x * dts + gvo.x + geo.x * h;
  const R gny = gxo.y * dts + gvo.y + geo.y * h;
  const R gnz = gxo.z * dts + gvo.z + geo.z * h;
  R gvx = gnx + geo.x * h, gvy = gny + geo.y * h, gvz = gnz + geo.z * h;
  R gax = gnx * dt, gay = gny * dt, gaz = gnz * dt;
  if (sv2 > (R)P.VL2) {  // v_l = v VL/|v|: J = (VL/|v|)(I - v v^T/|v|^2)
    const R sv = (R)P.VL / std::sqrt(sv2);
    const R d = (vel.x * gvx + vel.y * gvy + vel.z * gvz) / sv2;
    gvx = (gvx - vel.x * d) * sv;
    gvy = (gvy - vel.y * d) * sv;
    gvz = (gvz - vel.z * d) * sv;
  }
  if (sa2 > (R)P.AL2) {
    const R sa = (R)P.AL / std::sqrt(sa2);
    const R d = (ax * gax + ay * gay + az * gaz) / sa2;
    gax = (gax - ax * d) * sa;
    gay = (gay - ay * d) * sa;
    gaz = (gaz - az * d) * sa;
  }
  gvel = V3<R>{gvx, gvy, gvz};
  gF = V3<R>{gax * m, gay * m, gaz * m};
  R px = gxo.x, py = gxo.y, pz = gxo.z;
  R ex = 0, ey = 0, ez = 0;
  if (w1) {
    nx = -slope; ny = R(1) - slope;
    const R dn = nx * gax + ny * gay;
    const R t = stiff * dn * ss;
    px += t * slope;
    py -= t;
    const R dd = damp * dn;
    ex -= dd * nx + fric * gax;
    ey -= dd * ny + fric * gay;
    ez -= fric * gaz;
  }
  if (w2) {
    const R dn = -gay;
    py += stiff * dn * ss;
    ex -= fric * gax;
    ey -= fric * gay - damp * dn;
    ez -= fric * gaz;
  }
  if (w3) {
    px -= gax * (wall * stiff * ss);
    if (P.bound_wall_force != 0.0f) {
      const R dwall = orc_cos(sarg) * (R(0.5) / R(400)) * (R)P.bound_wall_force;
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
  gx = V3<R>{px, py, pz};
  ge = V3<R>{ex, ey, ez};
}

// ---------------------------------------------------------- force (reverse)
template <class R>
struct Coef {
  R vterm, stiff, lap;
};

template <class R>
void force_bwd_one(const FParams_t& P, const Coef<R>& K, const TapeStep<R>& T, const std::vector<V3<R>>& gF,
                   int i, V3<R>& gx, V3<R>& ge, R& grho, R& gk, R& gvisc) {
  gx = ge = V3<R>{0, 0, 0};
  grho = gk = gvisc = 0;
  const int gc = T.gcell[i];
  if (gc == FX_GRID_UNDEF) return;
  const int base = gc - (P.gridRes.z + 1) * P.gridRes.x - 1;
  const int srch = P.gridSrch;
  const R rd2 = P.rd2, d2 = P.d2, h = P.psmoothradius, rest = P.prest_dens;
  const R kp = ((R)P.sim_scale * R(-0.5)) * (R)P.spikykern;
  const V3<R> p = T.pos[i], vi = T.veval[i], gi = gF[i];
  const R di = T.press[i];
  const R pi = (di - rest) * K.stiff;
  const R idi = R(1) / di;
  const R qi = di - rest;
  R gpx = 0, gpy = 0, gpz = 0, gex = 0, gey = 0, gez = 0, gr = 0, ak = 0, av = 0;
  for (int r = 0; r < srch * srch; r++) {
    const int c0 = base + P.gridAdj[r * srch];
    const int c1 = c0 + srch - 1;
    const int jb = (int)T.off[c0], je = (int)(T.off[c1] + T.cnt[c1]);
    for (int j = jb; j < je; j++) {
      const V3<R> q = T.pos[j];
      const R dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
      const R dsq = std::fma(dz, dz, std::fma(dy, dy, dx * dx));
      if (!(dsq < rd2 && dsq > R(0))) continue;
      const R rr = std::sqrt(dsq * d2);
      const R c = h - rr;
      const R dj = T.press[j];
      const R pj = (dj - rest) * K.stiff;
      const R inv = R(1) / (rr * (di * dj));
      const R invd = inv * rr;          // 1 / (rho_i rho_j)
      const R ir = inv * (di * dj);     // 1 / r
      const R kc2i = (kp * (c * c)) * inv;
      const R pp = pi + pj;
      const R A = kc2i * pp;
      const R B = (K.vterm * c) * invd;
      const V3<R> gj = gF[j], vj = T.veval[j];
      const R lx = gi.x - gj.x, ly = gi.y - gj.y, lz = gi.z - gj.z;
      const R wx = vj.x - vi.x, wy = vj.y - vi.y, wz = vj.z - vi.z;
      const R su = std::fma(dz, lz, std::fma(dy, ly, dx * lx));
      const R sv = std::fma(wz, lz, std::fma(wy, ly, wx * lx));
      // (d2/r) (A_r su + B_r sv), A_r = -kp pp c (2r + c) invd / r^2, B_r = -vterm invd
      const R Ar = -(kp * pp) * (c * (rr + rr + c)) * (invd * (ir * ir));
      const R Br = -K.vterm * invd;
      const R g = (d2 * ir) * std::fma(Ar, su, Br * sv);
      gpx = std::fma(A, lx, std::fma(g, dx, gpx));
      gpy = std::fma(A, ly, std::fma(g, dy, gpy));
      gpz = std::fma(A, lz, std::fma(g, dz, gpz));
      gex = std::fma(-B, lx, gex);
      gey = std::fma(-B, ly, gey);
      gez = std::fma(-B, lz, gez);
      gr = std::fma(kc2i * K.stiff, su, gr) - std::fma(A, su, B * sv) * idi;
      const R ui = std::fma(dz, gi.z, std::fma(dy, gi.y, dx * gi.x));
      const R wi = std::fma(wz, gi.z, std::fma(wy, gi.y, wx * gi.x));
      ak = std::fma(kc2i * (qi + (dj - rest)), ui, ak);
      av = std::fma((K.lap * c) * invd, wi, av);
    }
  }
  gx = V3<R>{gpx, gpy, gpz};
  ge = V3<R>{gex, gey, gez};
  grho = gr;
  gk = ak;
  gvisc = av;
}

// -------------------------------------------------------- density (reverse)
template <class R>
V3<R> density_bwd_one(const FParams_t& P, const TapeStep<R>& T, const std::vector<R>& grho, int i) {
  const int gc = T.gcell[i];
  if (gc == FX_GRID_UNDEF) return V3<R>{0, 0, 0};
  const int base = gc - (P.gridRes.z + 1) * P.gridRes.x - 1;
  const int srch = P.gridSrch;
  const R rd2 = P.rd2, d2 = P.d2;
  const R k6 = ((R(-6) * (R)P.pmass) * (R)P.poly6kern) * d2;
  const V3<R> p = T.pos[i];
  const R gi = grho[i];
  R gpx = 0, gpy = 0, gpz = 0;
  for (int r = 0; r < srch * srch; r++) {
    const int c0 = base + P.gridAdj[r * srch];
    const int c1 = c0 + srch - 1;
    const int jb = (int)T.off[c0], je = (int)(T.off[c1] + T.cnt[c1]);
    for (int j = jb; j < je; j++) {
      const V3<R> q = T.pos[j];
      const R dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
      const R dsq = std::fma(dz, dz, std::fma(dy, dy, dx * dx));
      if (!(dsq < rd2 && dsq > R(0))) continue;
      const R t = (rd2 - dsq) * d2;
      const R cf = ((gi + grho[j]) * (t * t)) * k6;
      gpx = std::fma(cf, dx, gpx);
      gpy = std::fma(cf, dy, gpy);
      gpz = std::fma(cf, dz, gpz);
    }
  }
  return V3<R>{gpx, gpy, gpz};
}

// One reverse step over tape record T; G holds the adjoint of the step's
// outputs on entry and of its inputs on return (id frame).
template <class R>
void backward_step(const FParams_t& P, const Coef<R>& K, const TapeStep<R>& T, Adjoint<R>& G, int threads) {
  const int n = (int)T.pos.size();
  const R dt = P.dt, ss = P.sim_scale;
  std::vector<V3<R>> gF(n), gxs(n), ges(n);
  std::vector<R> grho(n), pk(n), pv(n);
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
  for (int s = 0; s < n; s++) {
    const uint32_t id = T.id[s];
    V3<R> gvel;
    advance_bwd_one(P, T, s, dt, ss, G.gx[id], G.gv[id], G.ge[id], gF[s], gxs[s], ges[s], gvel);
    G.gv[id] = gvel;
  }
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads) if (threads > 1)
  for (int i = 0; i < n; i++) {
    V3<R> gx, ge;
    force_bwd_one(P, K, T, gF, i, gx, ge, grho[i], pk[i], pv[i]);
    gxs[i] = V3<R>{gxs[i].x + gx.x, gxs[i].y + gx.y, gxs[i].z + gx.z};
    ges[i] = V3<R>{ges[i].x + ge.x, ges[i].y + ge.y, ges[i].z + ge.z};
  }
  for (int i = 0; i < n; i++) {
    G.gk += (double)pk[i];
    G.gvisc += (double)pv[i];
  }
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads) if (threads > 1)
  for (int i = 0; i < n; i++) {
    const V3<R> g = density_bwd_one(P, T, grho, i);
    const uint32_t id = T.id[i];
    G.gx[id] = V3<R>{gxs[i].x + g.x, gxs[i].y + g.y, gxs[i].z + g.z};
    G.ge[id] = ges[i];
  }
}

template <class R>
Coef<R> coef_of(const FParams_t& P, const State<R>& S) {
  Coef<R> K;
  K.vterm = S.coef_set ? S.vterm : (R)P.vterm;
  K.stiff = S.coef_set ? S.stiff : (R)P.pintstiff;
  K.lap = P.lapkern;
  return K;
}

// Loss of an nsteps rollout from S against target (by id), and its gradient
// with respect to the initial velocities, evaluation velocities (by id),
// viscosity and internal stiffness.
template <class R>
double loss_and_grad(const FParams_t& P, State<R>& S, int nsteps, const R* target, Adjoint<R>& G, int threads) {
  std::vector<TapeStep<R>> tape;
  const Coef<R> K = coef_of(P, S);
  rollout(P, S, nsteps, &tape, threads);
  const double L = loss_seed(S, target, G);
  for (int t = nsteps - 1; t >= 0; t--) backward_step(P, K, tape[t], G, threads);
  return L;
}

// FP64 adjoint over an FP32 forward: the FP32 rollout's tape promoted to
// double.  Separates the adjoint's own rounding from the divergence of the
// FP32 trajectory from the FP64 one.
template <class Src, class Dst>
void promote(const TapeStep<Src>& a, TapeStep<Dst>& b) {
  auto v3 = [](const std::vector<V3<Src>>& x, std::vector<V3<Dst>>& y) {
    y.resize(x.size());
    for (size_t i = 0; i < x.size(); i++) y[i] = V3<Dst>{(Dst)x[i].x, (Dst)x[i].y, (Dst)x[i].z};
  };
  v3(a.pos, b.pos);
  v3(a.vel, b.vel);
  v3(a.veval, b.veval);
  v3(a.force, b.force);
  b.press.assign(a.press.begin(), a.press.end());
  b.id = a.id;
  b.gcell = a.gcell;
  b.cnt = a.cnt;
  b.off = a.off;
  b.time = (Dst)a.time;
}

inline double loss_and_grad_mixed(const FParams_t& P, State<float>& S, int nsteps, const float* target,
                                  Adjoint<double>& G, int threads) {
  std::vector<TapeStep<float>> tape;
  rollout(P, S, nsteps, &tape, threads);
  State<double> F;
  F.n = S.n;
  F.G = S.G;
  F.pos.resize(S.n);
  for (int i = 0; i < S.n; i++) F.pos[i] = V3<double>{S.pos[i].x, S.pos[i].y, S.pos[i].z};
  F.id = S.id;
  std::vector<double> tgt(3 * (size_t)S.n);
  for (size_t i = 0; i < tgt.size(); i++) tgt[i] = target[i];
  const double L = loss_seed(F, tgt.data(), G);
  Coef<double> K;
  K.vterm = (double)P.vterm;
  K.stiff = (double)P.pintstiff;
  K.lap = (double)P.lapkern;
  TapeStep<double> T;
  for (int t = nsteps - 1; t >= 0; t--) {
    promote(tape[t], T);
    backward_step(P, K, T, G, threads);
  }
  return L;
}

}  // namespace oracle
