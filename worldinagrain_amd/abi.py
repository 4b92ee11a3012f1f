"""ctypes mirror of include/wig.h: the structs and constants of the C ABI."""
import ctypes as ct

WIG_OK = 0
WIG_WARN_OVERFLOW = 1
WIG_NOT_READY = 2
WIG_E_INVALID = -1
WIG_E_NOMEM = -2
WIG_E_DEVICE = -3
WIG_E_NOT_BUILT = -4

WIG_TOL_ABS_MAX = 0
WIG_TOL_REL_L2 = 1
WIG_STORE_F32, WIG_STORE_F16, WIG_STORE_BF16 = 0, 1, 2
WIG_MAX_CYCLES = 256

PART_CELLS, PART_PARTICLES0, PART_PARTICLES1, PART_COUNTERS, PART_FIELD, PART_BLOCKS, PART_FLUIDCG, PART_SOLVE = range(8)
PART_AIR, PART_AIR_SOURCE, PART_OCCUPANTS, PART_GRAVITY = 8, 9, 10, 11
PART_MOTION, PART_SWEPT, PART_PUSHES = 12, 13, 14
PART_COUNT = 15
LIQUID_PARTS = 8          # the liquid's own parts (a snapshot's); the rest are the caller's (the air's, the people's, the bodies')
PART_NAMES = ("cells", "particles0", "particles1", "counters", "field", "blocks", "fluidCg", "solve", "air0",
              "airSource", "occupants", "gravity", "bodyMotion", "bodySwept", "bodyPushes")
# the bodies (js/shared.js:299-333; common.wgsl:353-387): the table's slots
# and words (BODY_WORD), their flags (BODY_FLAG), the liquid's slots after the
# contacts, the push sets
MAX_BODIES = 1024
BODY_TABLE_WORDS = 116
BODY_CONTACTS, BODY_CONTACT_WORDS, BODY_LIQUID_WORDS = 16384, 12, 4
BODY_PUSH_WORDS = 36
MOTION_WORDS = MAX_BODIES * (1 + BODY_TABLE_WORDS) + BODY_CONTACTS * BODY_CONTACT_WORDS + MAX_BODIES * BODY_LIQUID_WORDS
BW_POS, BW_VEL, BW_SPIN, BW_MASS, BW_INERTIA, BW_CELLS, BW_FLAGS, BW_FRAME = 0, 3, 5, 6, 7, 13, 18, 19
BW_SUBCELL, BW_CARGO, BW_CARGO_INERTIA, BW_MOBILITY, BW_MOBILITY_AT = 25, 105, 106, 108, 114
BF_ALIVE, BF_NEW, BF_HELD, BF_WET = 1, 2, 32, 64
FL_LOOSE = 8
RIGID_FIX = 16777216.0
EXACT_COARSE = 16777216.0
LIQUID_WET = 0x80000000
CLAIM_PERIOD = 0xFFFFF
PUSH_P0, PUSH_X, PUSH_F, PUSH_SURFACE, PUSH_DRAG = 0, 2, 3, 4, 5
# an element's columns (WigElementProps): viscosity, viscosity_e, viscosity_t,
# tension, wetting, breaks
PROP_COLUMNS = 6


def liquid_slot(b, w):
    """Word w of body b's liquid slot in the motion part (fluid.wgsl:1226-1228)."""
    return MAX_BODIES * (1 + BODY_TABLE_WORDS) + BODY_CONTACTS * BODY_CONTACT_WORDS + b * BODY_LIQUID_WORDS + w


def body_word(b, k):
    """Word k of body b's slot in the motion part's table (fluid.wgsl:1211-1213)."""
    return MAX_BODIES + b * BODY_TABLE_WORDS + k


def claim_step(frame):
    """common.wgsl:384-386."""
    return (frame % CLAIM_PERIOD) + 1


def swept_key(frame, b, seam=False):
    """A swept word for body b stamped at `frame` (common.wgsl:368-370; bit 0 a seam)."""
    return ((claim_step(frame) << 12) | ((2047 - b) << 1) | (1 if seam else 0)) & 0xFFFFFFFF


def element_props(props):
    """props as the C ABI takes them: (256, PROP_COLUMNS) f32, short rows padded with 0."""
    import numpy as np
    p = np.asarray(props, dtype=np.float32)
    assert p.ndim == 2 and p.shape[0] == 256 and p.shape[1] <= PROP_COLUMNS
    out = np.zeros((256, PROP_COLUMNS), dtype=np.float32)
    out[:, :p.shape[1]] = p
    return out
# the gravity pass's field: a vec4 (pull x, y, potential, down) a light cell
# of GRAVITY_BLOCK x GRAVITY_BLOCK cells (pull.wgsl:1-44)
GRAVITY_BLOCK = 2
DOWN_WEIGHTLESS = 8
# occupantWord (common.wgsl:238-254): slot + 1 in the low byte, velocity x 128
# in two 12-bit fields (x at bit 8, y at bit 20)
OCCUPANT_SCALE = 128.0
# airSource's words an air cell (airsource.wgsl:18-25), each an exact sum of two words
SOURCE_WORDS = 16
SOURCE_RIGHT, SOURCE_DOWN, SOURCE_AMOUNT, SOURCE_VAPOUR, SOURCE_MIST, SOURCE_LOST, SOURCE_HEAT = range(7)
AIR_BLOCKED, AIR_SOLID = -10000.0, -20000.0

STAT_ACTIVE, STAT_RETIRED, STAT_VISCOUS, STAT_PARTICLES, STAT_RESIDUAL = 0, 1, 2, 3, 4
STAT_CYCLES, STAT_MOST_CYCLES = 7, 8
STATS = 9


class WigParams(ct.Structure):
    """`Params`, 80 bytes (shaders/common.wgsl:5-42; js/sim.js:1972-2001)."""
    _fields_ = [
        ("size", ct.c_uint32 * 2), ("frame", ct.c_uint32), ("seed", ct.c_uint32),
        ("offset", ct.c_uint32 * 2), ("supportMax", ct.c_float), ("dx", ct.c_float),
        ("day", ct.c_float), ("dayRate", ct.c_float), ("holePull", ct.c_float), ("massPull", ct.c_float),
        ("earthG", ct.c_float), ("dt", ct.c_float), ("dayDt", ct.c_float), ("clockSpan", ct.c_float),
        ("heatStages", ct.c_uint32), ("heatLimit", ct.c_float), ("holding", ct.c_uint32), ("pad", ct.c_float),
    ]


class WigStepCtl(ct.Structure):
    _fields_ = [
        ("full", ct.c_uint32), ("viscous", ct.c_uint32), ("surface", ct.c_uint32), ("bodies_run", ct.c_uint32),
        ("fluid_cycles", ct.c_uint32), ("shifted_first", ct.c_uint32), ("tol_mode", ct.c_uint32), ("tol", ct.c_float),
        ("pre_relax", ct.c_uint32), ("post_relax", ct.c_uint32), ("sweep_pairs", ct.c_uint32),
        ("coarse_sweeps", ct.c_uint32), ("first_chunk", ct.c_uint32), ("chunk", ct.c_uint32),
        ("level_pairs", ct.c_uint32), ("vcycle_storage", ct.c_uint32), ("residual_every", ct.c_uint32),
        ("air", ct.c_uint32), ("surface_drag", ct.c_float), ("occupants", ct.c_uint32),
        ("gravity", ct.c_uint32), ("bodies", ct.c_uint32),
    ]


class WigLayoutInfo(ct.Structure):
    _fields_ = [
        ("width", ct.c_uint32), ("height", ct.c_uint32), ("cells", ct.c_uint32), ("faces", ct.c_uint32),
        ("blocks", ct.c_uint32), ("blocks_wide", ct.c_uint32), ("capacity", ct.c_uint32), ("levels", ct.c_uint32),
        ("level_width", ct.c_uint32 * 9), ("level_height", ct.c_uint32 * 9), ("part_bytes", ct.c_uint64 * PART_COUNT),
    ]


class WigSolveStats(ct.Structure):
    _fields_ = [
        ("rz", ct.c_double), ("alpha", ct.c_double), ("beta", ct.c_double), ("pq", ct.c_double),
        ("rr", ct.c_double), ("bb", ct.c_double), ("cycles", ct.c_uint32), ("converged", ct.c_uint32),
        ("vcycle_scale", ct.c_float), ("pad", ct.c_uint32),
    ]


class WigCheck64(ct.Structure):
    _fields_ = [("iterations", ct.c_uint32), ("pad", ct.c_uint32), ("relres64", ct.c_double), ("err", ct.c_double),
                ("relres", ct.c_double)]


assert ct.sizeof(WigParams) == 80
assert ct.sizeof(WigSolveStats) == 64


class WigRunTotals(ct.Structure):
    _fields_ = [("steps", ct.c_uint64), ("cycles", ct.c_uint64), ("particles", ct.c_uint64),
                ("active_cells", ct.c_uint64)]


def occupant_word(slot, vx, vy):
    """common.wgsl's occupantWord (:240-243): a person's slot and velocity
    (cells a step) as the occupants part's word."""
    import numpy as np
    q = np.clip(np.rint(np.array([vx, vy], dtype=np.float32) * np.float32(OCCUPANT_SCALE)), -2047.0, 2047.0).astype(np.int32)
    return int((slot + 1) | ((int(q[0]) & 0xFFF) << 8) | ((int(q[1]) & 0xFFF) << 20))
