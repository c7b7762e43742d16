"""ctypes mirrors of the plain-C types in include/fluidsx_params.h and include/fluidsx.h.

FParams_t follows fluids5.0/src/fluid.h:45-82 member for member (688 bytes);
fx_scene is the data form of what the reference hard-codes in
Particles::SetupDefaultParams / SetupExampleParams (fluids5.0/src/particles.cpp:122-263).
"""
import ctypes as C

GRID_UNDEF = 2147483647  # fluid.h:12

# particle buffers (fluid.h:15-22) plus FID
FPOS, FVEL, FCLR, FVEVAL, FFORCE, FPRESS, FGCELL, FGNDX, FID = range(9)
# grid buffers (fluid.h:25-31)
AGRID, AGRIDCNT, AGRIDOFF, AAUXARRAY1, AAUXSCAN1, AAUXARRAY2, AAUXSCAN2 = range(7)
ANBRBITS, AFREC, ATAILFLAG, ATAILSCAN, ATAILAUX1, ATAILAUXS1, ATAILAUX2, ATAILAUXS2, ASLABSTART = range(7, 16)
FX_SLABS = 16
FX_MEMBER_BITS = 28  # AGRID member entries: index in the low bits, x slab above
NBR_WORDS = 256  # FX_NBR_WORDS: bitmask words per particle (8192 candidates)

# stages of Particles::Run (particles.cpp:382-394)
STAGE_INSERT, STAGE_PREFIX, STAGE_SORT, STAGE_PRESSURE, STAGE_FORCE, STAGE_ADVANCE = range(6)
NUM_STAGES = 6
STAGE_NAMES = ("insert", "prefix", "sort", "pressure", "force", "advance")

FX_OK, FX_EINVAL, FX_EHIP, FX_ESTATE, FX_ERANGE, FX_EMISMATCH = 0, -1, -2, -3, -4, -5

# element layout of each particle buffer: (numpy dtype, components)
PBUF_LAYOUT = {
    FPOS: ("float32", 3), FVEL: ("float32", 3), FVEVAL: ("float32", 3), FFORCE: ("float32", 3),
    FCLR: ("uint32", 1), FPRESS: ("float32", 1), FGCELL: ("int32", 1), FGNDX: ("int32", 1),
    FID: ("uint32", 1),
}


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]

    def tolist(self):
        return [self.x, self.y, self.z]


class Int3(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]

    def tolist(self):
        return [self.x, self.y, self.z]


def f3(v):
    return Float3(*[float(a) for a in v])


class FParams(C.Structure):
    _fields_ = [
        ("mode", C.c_int), ("example", C.c_int),
        ("time", C.c_float), ("dt", C.c_float), ("sim_scale", C.c_float),
        ("capture", C.c_ubyte),
        ("numThreads", C.c_int), ("numBlocks", C.c_int),
        ("gridThreads", C.c_int), ("gridBlocks", C.c_int),
        ("szPnts", C.c_int), ("szHash", C.c_int), ("szGrid", C.c_int),
        ("stride", C.c_int), ("pnum", C.c_int), ("chk", C.c_int),
        ("pdist", C.c_float), ("pmass", C.c_float), ("prest_dens", C.c_float), ("pintstiff", C.c_float),
        ("pradius", C.c_float), ("psmoothradius", C.c_float), ("pspacing", C.c_float), ("r2", C.c_float),
        ("pvisc", C.c_float),
        ("AL", C.c_float), ("AL2", C.c_float), ("VL", C.c_float), ("VL2", C.c_float),
        ("d2", C.c_float), ("rd2", C.c_float), ("vterm", C.c_float),
        ("poly6kern", C.c_float), ("spikykern", C.c_float), ("lapkern", C.c_float), ("gausskern", C.c_float),
        ("grid_size", C.c_float), ("grid_density", C.c_float),
        ("gridSize", Float3), ("gridDelta", Float3), ("gridMin", Float3), ("gridMax", Float3),
        ("gridRes", Int3), ("gridScanMax", Int3),
        ("gridSrch", C.c_int), ("gridTotal", C.c_int), ("gridAdjCnt", C.c_int), ("gridActive", C.c_int),
        ("gridAdj", C.c_int * 64),
        ("bound_slope", C.c_float), ("bound_stiff", C.c_float), ("bound_friction", C.c_float),
        ("bound_damp", C.c_float),
        ("bound_wall_force", C.c_float), ("bound_wall_freq", C.c_float),
        ("bound_min", Float3), ("bound_max", Float3),
        ("init_min", Float3), ("init_max", Float3),
        ("grav_dir", Float3), ("grav_pos", Float3), ("gravity", Float3),
        ("grav_amt", C.c_float),
        ("emit_rate", Float3), ("emit_pos", Float3), ("emit_ang", Float3), ("emit_dang", Float3),
        ("emit_spread", Float3),
        ("brickRes", Int3),
        ("pemit", C.c_int),
        ("pmem", C.c_int),
    ]


assert C.sizeof(FParams) == 688, C.sizeof(FParams)


class Scene(C.Structure):
    _fields_ = [
        ("num_points", C.c_int32), ("reserved0", C.c_int32),
        ("seed", C.c_uint64),
        ("dt", C.c_float), ("sim_scale", C.c_float), ("smooth_radius", C.c_float), ("mass", C.c_float),
        ("rest_density", C.c_float), ("visc", C.c_float), ("int_stiff", C.c_float),
        ("bound_stiff", C.c_float), ("bound_damp", C.c_float), ("accel_limit", C.c_float),
        ("vel_limit", C.c_float),
        ("bound_slope", C.c_float), ("bound_friction", C.c_float), ("bound_wall_force", C.c_float),
        ("bound_wall_freq", C.c_float),
        ("pradius", C.c_float),
        ("grid_density", C.c_float),
        ("grav_amt", C.c_float),
        ("grav_dir", Float3),
        ("bound_min", Float3), ("bound_max", Float3), ("init_min", Float3), ("init_max", Float3),
        ("spacing", C.c_float),
        ("jitter", C.c_float),
    ]

    def copy(self):
        s = Scene()
        C.memmove(C.byref(s), C.byref(self), C.sizeof(Scene))
        return s


def params_copy(p):
    q = FParams()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(FParams))
    return q
