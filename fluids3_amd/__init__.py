"""fluids3_amd -- the Fluids 5.0 SPH step as gfx950 HIP kernels behind the reference's
Particles / DataX interface, with a C ABI (include/fluidsx.h, libfluidsx.so)."""
from . import abi, scenes  # noqa: F401
from .particles import Particles, FluidsError, derive_params, fill_points, scene_spacing  # noqa: F401
