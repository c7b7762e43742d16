"""Dam-break scenes of the benchmark configurations, as fx_scene structs.

Physical parameters (all configurations): dt=0.003, sim_scale=0.004, smoothing
radius h=0.01, mass=0.00020543, rest density 600, viscosity 0.35, internal
stiffness 1.0, boundary stiffness 10000, boundary damping 256, accel limit 150,
velocity limit 3, gravity (0,-9.8,0).  The grid cell edge is h/sim_scale,
i.e. grid_density = 1.5 in Particles::RebuildAccelGrid (particles.cpp:551).
The boundary contact radius pradius is h, the reference's convention
(particles.cpp:165-167).

Initial state: cubic lattice filled y-layer by y-layer (x fastest, then z)
from the block's minimum corner, with seeded uniform jitter of +-0.1 spacing
per axis and zero velocity.  The spacing is the rest spacing
(mass/rest_density)^(1/3)/sim_scale unless the block cannot hold N particles
at that spacing; then it is the largest spacing whose lattice holds N
(fx_scene_spacing).
"""
import math

from .abi import Scene, f3

SEED = 20240531


def dam_break(n, box_min, box_max, block_min, block_max, seed=SEED, visc=0.35, int_stiff=1.0):
    s = Scene()
    s.num_points = int(n)
    s.seed = int(seed)
    s.dt = 0.003
    s.sim_scale = 0.004
    s.smooth_radius = 0.01
    s.mass = 0.00020543
    s.rest_density = 600.0
    s.visc = float(visc)
    s.int_stiff = float(int_stiff)
    s.bound_stiff = 10000.0
    s.bound_damp = 256.0
    s.accel_limit = 150.0
    s.vel_limit = 3.0
    s.bound_slope = 0.0
    s.bound_friction = 0.0
    s.bound_wall_force = 0.0
    s.bound_wall_freq = 0.0
    s.pradius = 0.01
    s.grid_density = 1.5
    s.grav_amt = 1.0
    s.grav_dir = f3((0.0, -9.8, 0.0))
    s.bound_min = f3(box_min)
    s.bound_max = f3(box_max)
    s.init_min = f3(block_min)
    s.init_max = f3(block_max)
    s.spacing = 0.0
    s.jitter = 0.1
    return s


def rest_spacing():
    return (0.00020543 / 600.0) ** (1.0 / 3.0) / 0.004


def dambreak_16k(n=16384, seed=SEED):
    """16,384 particles; box (-30,0,-30)..(30,60,30), block (-25,5,-25)..(0,45,25)."""
    return dam_break(n, (-30, 0, -30), (30, 60, 30), (-25, 5, -25), (0, 45, 25), seed=seed)


def dambreak_4m(n=4_000_000, seed=SEED):
    """4,000,000 particles; box (-120,0,-60)..(120,120,60), block = left 40% in x.

    At the stated size the block holds the particles at 0.40x rest spacing
    (about 15x rest density): the literal reading of the configuration.
    """
    return dam_break(n, (-120, 0, -60), (120, 120, 60), (-120, 0, -60), (-24, 120, 60), seed=seed)


def scaled_dam_break(n, seed=SEED, visc=0.35, int_stiff=1.0, fill=0.4):
    """Box of aspect 2:1:1 whose left `fill` fraction in x holds n particles at rest spacing."""
    s = rest_spacing()
    # block = fill * Lx * L * L with Lx = 2L  ->  2 fill L^3 = n s^3 (3 % slack for the lattice's ceil)
    L = (n * s ** 3 * 1.03 / (2.0 * fill)) ** (1.0 / 3.0)
    lx = 2.0 * L
    bmin = (-L, 0.0, -L / 2)
    bmax = (L, L, L / 2)
    blk_max = (-L + fill * lx, L, L / 2)
    return dam_break(n, bmin, bmax, bmin, blk_max, seed=seed, visc=visc, int_stiff=int_stiff)


def dambreak_4m_rest(n=4_000_000, seed=SEED):
    """4,000,000 particles at rest spacing: the dambreak_4m box scaled up until its left 40% holds them."""
    return scaled_dam_break(n, seed=seed)


def batch_scenes(count=64, lo=8192, hi=262144, seed=SEED, visc_range=(0.2, 0.5)):
    """Independent scenes with log-uniform particle counts in [lo, hi] and uniform viscosity."""
    import numpy as np

    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(math.log(lo), math.log(hi), size=count)).astype(np.int64)
    viscs = rng.uniform(visc_range[0], visc_range[1], size=count)
    return [scaled_dam_break(int(n), seed=seed + 1 + k, visc=float(v)) for k, (n, v) in enumerate(zip(ns, viscs))]


def dambreak_262k(n=262144, seed=SEED):
    """Differentiable dam break: 262,144 particles at rest spacing, fill 0.4."""
    return scaled_dam_break(n, seed=seed)


def batch32_scenes(count=32, seed=SEED):
    """Batched adjoint: counts log-uniform in [16,384, 1,048,576], viscosity uniform in [0.2, 0.5]."""
    return batch_scenes(count, lo=16384, hi=1048576, seed=seed)


def ref_beach(n=4_000_000, seed=SEED):
    """The reference's own default scene, the one its published 4M-particle rate
    is measured on: Fluids 5.0 example 2, "large beach front"
    (Particles::SetupDefaultParams and SetupExampleParams,
    fluids5.0/src/particles.cpp:158-193, 246-252; grid_density 2.0 from
    Particles::Initialize, :86): box (0,0,0)..(500,200,500), block
    (120,60,0)..(500,195,500), h = pradius = 0.015, sim_scale 0.006, rest density
    200, mass 0.00020543, viscosity 0.1, internal stiffness 1.5, boundary
    stiffness 50000 and damping 150, limits 150 / 50, gravity (0,-4.8,0), ground
    slope 0.12, moving X- wall of amplitude 100 at frequency 2, dt 0.004.  The
    lattice spacing is AddPointsInVolume's pdist * 0.87 / sim_scale
    (particles.cpp:321-325) and the jitter its uniform [0, spacing) per axis
    (here: the lattice shifted by half a spacing, +-0.5 spacing seeded)."""
    s = dam_break(n, (0, 0, 0), (500, 200, 500), (120, 60, 0), (500, 195, 500), seed=seed, visc=0.1, int_stiff=1.5)
    s.dt = 0.004
    s.sim_scale = 0.006
    s.smooth_radius = 0.015
    s.pradius = 0.015
    s.mass = 0.00020543
    s.rest_density = 200.0
    s.bound_stiff = 50000.0
    s.bound_damp = 150.0
    s.accel_limit = 150.0
    s.vel_limit = 50.0
    s.bound_slope = 0.12
    s.bound_wall_force = 100.0
    s.bound_wall_freq = 2.0
    s.grid_density = 2.0
    s.grav_amt = 1.0
    s.grav_dir = f3((0.0, -4.8, 0.0))
    import numpy as np

    pdist = np.float32(np.float32(s.mass) / np.float32(s.rest_density)) ** np.float32(1.0 / 3.0)
    sp = float(np.float32(pdist * np.float32(0.87)) / np.float32(s.sim_scale))
    s.spacing = sp
    s.jitter = 0.5
    s.init_min = f3((120 + 0.5 * sp, 60 + 0.5 * sp, 0 + 0.5 * sp))
    s.init_max = f3((500 + 0.5 * sp, 195 + 0.5 * sp, 500 + 0.5 * sp))
    return s


PRESETS = {"dambreak_16k": dambreak_16k, "dambreak_4m": dambreak_4m, "dambreak_4m_rest": dambreak_4m_rest,
           "dambreak_262k": dambreak_262k, "ref_beach": ref_beach}
