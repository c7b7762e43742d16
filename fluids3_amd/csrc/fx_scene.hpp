// fx_scene.hpp -- host-only parameter, grid and lattice set-up.
#pragma once

#include "fluidsx.h"

namespace fx {

float jitter_uniform(uint64_t seed, uint64_t idx, int axis);
float scene_spacing(const fx_scene& sc);
int scene_derive(const fx_scene& sc, FParams_t& P);
int rebuild_grid(FParams_t& P);
void update_params(FParams_t& P);
void set_geometry(FParams_t& P, int threads);
int scene_fill(const fx_scene& sc, float* pos, uint32_t* clr, int32_t n);

}  // namespace fx
