// scene.cpp -- host-only set-up: parameters, grid and initial lattice.
//
// Follows the reference's float32 arithmetic exactly where it fixes integer
// outcomes (grid resolution, cell of each particle):
//   Particles::RebuildAccelGrid  fluids5.0/src/particles.cpp:548-626
//   Particles::UpdateParams      fluids5.0/src/particles.cpp:657-679
//   Particles::AddPointsInVolume fluids5.0/src/particles.cpp:318-379
#include <cmath>
#include <cstdint>
#include <cstring>

#include "fx_common.hpp"
#include "fx_scene.hpp"

namespace fx {

static inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

float jitter_uniform(uint64_t seed, uint64_t idx, int axis) {
  // counter-based: the value of particle idx, axis a depends on (seed, idx, a) only
  const uint64_t z = splitmix64(seed ^ splitmix64(idx * 3ull + (uint64_t)axis));
  return (float)(z >> 40) * (1.0f / 16777216.0f);  // [0, 1)
}

static uint64_t lattice_capacity(const fx_scene& sc, float s) {
  const float lx = sc.init_max.x - sc.init_min.x;
  const float ly = sc.init_max.y - sc.init_min.y;
  const float lz = sc.init_max.z - sc.init_min.z;
  if (!(s > 0.0f) || lx < 0 || ly < 0 || lz < 0) return 0;
  const uint64_t cx = (uint64_t)std::ceil(lx / s);
  const uint64_t cz = (uint64_t)std::ceil(lz / s);
  const uint64_t ny = (uint64_t)std::floor(ly / s) + 1;
  return (cx ? cx : 1) * (cz ? cz : 1) * ny;
}

float scene_spacing(const fx_scene& sc) {
  if (sc.spacing > 0.0f) return sc.spacing;
  // rest spacing (m/rho0)^(1/3) / ss, as in particles.cpp:323 without the
  // 0.87 packing factor
  float s = std::pow(sc.mass / sc.rest_density, 1.0f / 3.0f) / sc.sim_scale;
  const uint64_t n = (uint64_t)sc.num_points;
  if (lattice_capacity(sc, s) >= n) return s;
  const double vol = (double)(sc.init_max.x - sc.init_min.x) * (sc.init_max.y - sc.init_min.y) *
                     (sc.init_max.z - sc.init_min.z);
  s = (float)std::cbrt(vol / (double)n);
  for (int it = 0; it < 100000 && lattice_capacity(sc, s) < n; it++) s *= 0.9999f;
  return s;
}

int scene_derive(const fx_scene& sc, FParams_t& P) {
  if (sc.num_points <= 0) return fail(FX_EINVAL, "fx_scene_derive: num_points must be > 0");
  if (!(sc.sim_scale > 0) || !(sc.smooth_radius > 0) || !(sc.grid_density > 0) || !(sc.mass > 0) ||
      !(sc.rest_density > 0))
    return fail(FX_EINVAL, "fx_scene_derive: sim_scale, smooth_radius, grid_density, mass and "
                           "rest_density must be > 0");
  memset(&P, 0, sizeof(P));
  // SetupDefaultParams / SetupExampleParams inputs (particles.cpp:158-206,246-252)
  P.time = 0.0f;
  P.dt = sc.dt;
  P.sim_scale = sc.sim_scale;
  P.pvisc = sc.visc;
  P.prest_dens = sc.rest_density;
  P.pmass = sc.mass;
  P.pradius = sc.pradius;
  P.psmoothradius = sc.smooth_radius;
  P.pintstiff = sc.int_stiff;
  P.bound_stiff = sc.bound_stiff;
  P.bound_damp = sc.bound_damp;
  P.AL = sc.accel_limit;
  P.VL = sc.vel_limit;
  P.bound_slope = sc.bound_slope;
  P.bound_friction = sc.bound_friction;
  P.bound_wall_force = sc.bound_wall_force;
  P.bound_wall_freq = sc.bound_wall_freq;
  P.grav_amt = sc.grav_amt;
  P.grav_dir = sc.grav_dir;
  P.grav_pos = fx_float3{0, 0, 0};
  P.bound_min = sc.bound_min;
  P.bound_max = sc.bound_max;
  P.init_min = sc.init_min;
  P.init_max = sc.init_max;
  P.grid_density = sc.grid_density;
  P.pnum = sc.num_points;
  P.pspacing = scene_spacing(sc);
  FX_TRY(rebuild_grid(P));
  update_params(P);
  set_geometry(P, 256);
  return FX_OK;
}

int rebuild_grid(FParams_t& P) {
  // particles.cpp:551-580, with the reference's double/float promotions
  P.grid_size = (float)(1.5 * (double)P.psmoothradius / (double)P.grid_density);
  const float goff = (float)(2.0 * (double)(P.grid_size / P.sim_scale));
  P.gridMin = fx_float3{P.bound_min.x - goff, P.bound_min.y - goff, P.bound_min.z - goff};
  P.gridMax = fx_float3{P.bound_max.x + goff, P.bound_max.y + goff, P.bound_max.z + goff};
  P.gridSize = fx_float3{P.gridMax.x - P.gridMin.x, P.gridMax.y - P.gridMin.y, P.gridMax.z - P.gridMin.z};
  const float wc = P.grid_size / P.sim_scale;
  P.gridRes.x = (int)std::ceil(P.gridSize.x / wc);
  P.gridRes.y = (int)std::ceil(P.gridSize.y / wc);
  P.gridRes.z = (int)std::ceil(P.gridSize.z / wc);
  P.gridSize.x = (float)P.gridRes.x * wc;
  P.gridSize.y = (float)P.gridRes.y * wc;
  P.gridSize.z = (float)P.gridRes.z * wc;
  P.gridDelta.x = (float)P.gridRes.x / P.gridSize.x;
  P.gridDelta.y = (float)P.gridRes.y / P.gridSize.y;
  P.gridDelta.z = (float)P.gridRes.z / P.gridSize.z;
  const double total = (double)P.gridRes.x * P.gridRes.y * P.gridRes.z;
  if (P.gridRes.x <= 0 || P.gridRes.y <= 0 || P.gridRes.z <= 0 || total >= 1073741824.0)
    return fail(FX_ERANGE, "grid of %d x %d x %d cells is empty or beyond 1024^3", P.gridRes.x,
                P.gridRes.y, P.gridRes.z);
  P.gridTotal = P.gridRes.x * P.gridRes.y * P.gridRes.z;
  P.gridSrch = (int)(std::floor(2.0f * (P.psmoothradius / P.sim_scale) / wc) + 1.0f);
  if (P.gridSrch < 2) P.gridSrch = 2;
  // particles.cpp:582-585 stops above 6, but gridAdj holds 64 entries
  // (fluid.h:66), so more than 4 cells per axis would overrun it.
  if (P.gridSrch > 4)
    return fail(FX_ERANGE, "neighbour search of %d cells per axis exceeds 4 (gridAdj has 64 entries)",
                P.gridSrch);
  P.gridAdjCnt = P.gridSrch * P.gridSrch * P.gridSrch;
  P.gridScanMax = fx_int3{P.gridRes.x - P.gridSrch, P.gridRes.y - P.gridSrch, P.gridRes.z - P.gridSrch};
  int cell = 0;  // particles.cpp:614-618
  for (int y = 0; y < P.gridSrch; y++)
    for (int z = 0; z < P.gridSrch; z++)
      for (int x = 0; x < P.gridSrch; x++) P.gridAdj[cell++] = (y * P.gridRes.z + z) * P.gridRes.x + x;
  return FX_OK;
}

void update_params(FParams_t& P) {
  // particles.cpp:659-673 (pow(float, float) is powf)
  P.gravity = fx_float3{P.grav_dir.x * P.grav_amt, P.grav_dir.y * P.grav_amt, P.grav_dir.z * P.grav_amt};
  P.AL2 = P.AL * P.AL;
  P.VL2 = P.VL * P.VL;
  const float sr = P.psmoothradius;
  P.r2 = sr * sr;
  P.pdist = std::pow(P.pmass / P.prest_dens, 1 / 3.0f);
  P.poly6kern = 315.0f / (64.0f * 3.141592f * std::pow(sr, 9.0f));
  P.spikykern = -45.0f / (3.141592f * std::pow(sr, 6.0f));
  P.lapkern = 45.0f / (3.141592f * std::pow(sr, 6.0f));
  P.gausskern = 1.0f / std::pow(3.141592f * 2.0f * sr * sr, 3.0f / 2.0f);
  P.d2 = P.sim_scale * P.sim_scale;
  P.rd2 = P.r2 / P.d2;
  P.vterm = P.lapkern * P.pvisc;
}

void set_geometry(FParams_t& P, int threads) {
  // computeNumBlocks (particles.cpp:26-30) for pnum particles, with blocks
  // of whole wave64s
  const int waves = (P.pnum + 63) / 64 * 64;
  P.numThreads = threads < waves ? threads : waves;
  P.numBlocks = P.numThreads == 0 ? 1 : (P.pnum + P.numThreads - 1) / P.numThreads;
  P.szPnts = P.numBlocks * P.numThreads;
}

int scene_fill(const fx_scene& sc, float* pos, uint32_t* clr, int32_t n) {
  // particles.cpp:333-378: layers bottom-up in y, x fastest then z, stop at n.
  // Lattice points are min + k*s on every axis; jitter is uniform in
  // +-jitter*s per axis from the counter-based RNG.
  if (n < 0) return fail(FX_EINVAL, "fx_scene_fill: n < 0");
  const float s = scene_spacing(sc);
  const fx_float3 mn = sc.init_min, mx = sc.init_max;
  if (!(s > 0.0f) || !std::isfinite(s)) return fail(FX_EINVAL, "fx_scene_fill: lattice spacing must be > 0");
  if (!std::isfinite(mn.x) || !std::isfinite(mn.y) || !std::isfinite(mn.z) || !std::isfinite(mx.x) ||
      !std::isfinite(mx.y) || !std::isfinite(mx.z))
    return fail(FX_EINVAL, "fx_scene_fill: the fill volume must be finite");
  const double ex = std::ceil(((double)mx.x - mn.x) / s), ez = std::ceil(((double)mx.z - mn.z) / s);
  if (ex * (ez > 1.0 ? ez : 1.0) > (double)INT32_MAX)
    return fail(FX_ERANGE, "fx_scene_fill: a layer of %.0f x %.0f lattice sites is too large", ex, ez);
  const int cx = (int)ex, cz = (int)ez;
  const int cntx = cx > 0 ? cx : 1;
  const int cnt = cntx * (cz > 0 ? cz : 1);
  const float dx = mx.x - mn.x, dz = mx.z - mn.z;
  const float amp = sc.jitter * s;
  int p = 0;
  for (int ky = 0; p < n; ky++) {
    const float y = mn.y + (float)ky * s;
    if (y > mx.y) break;
    for (int xz = 0; xz < cnt && p < n; xz++, p++) {
      const float x = mn.x + (float)(xz % cntx) * s;
      const float z = mn.z + (float)(xz / cntx) * s;
      pos[3 * p + 0] = x + (2.0f * jitter_uniform(sc.seed, (uint64_t)p, 0) - 1.0f) * amp;
      pos[3 * p + 1] = y + (2.0f * jitter_uniform(sc.seed, (uint64_t)p, 1) - 1.0f) * amp;
      pos[3 * p + 2] = z + (2.0f * jitter_uniform(sc.seed, (uint64_t)p, 2) - 1.0f) * amp;
      if (clr != nullptr) {
        // particles.cpp:364-370 and COLORA (helpers/common_defs.h:128)
        const float px = dx != 0 ? (x - mn.x) / dx : 0.0f, pz = dz != 0 ? (z - mn.z) / dz : 0.0f;
        float r = 0.0f * 0.7f + 0.2f, g = px * 0.7f + 0.2f, b = (px + pz) * 0.5f * 0.7f + 0.2f;
        r = r < 0 ? 0 : (r > 1 ? 1 : r);
        g = g < 0 ? 0 : (g > 1 ? 1 : g);
        b = b < 0 ? 0 : (b > 1 ? 1 : b);
        clr[p] = ((uint32_t)(1.0f * 255.0f) << 24) | ((uint32_t)(b * 255.0f) << 16) |
                 ((uint32_t)(g * 255.0f) << 8) | (uint32_t)(r * 255.0f);
      }
    }
  }
  return p;
}

}  // namespace fx
