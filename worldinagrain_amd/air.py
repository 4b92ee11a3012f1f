"""The air's projection on the GPU (include/wig.h's wig_air_*): the MG-PCG
that solves for the air's dynamic pressure each step, before the liquid
step (shaders/air.wgsl:869-1536), on the reference's buffers.  There is no
CPU fallback: without libwig.so or a device this raises.

`AirProjection` keeps what the reference's host keeps for it: the
iterations to encode a step, `airCycles` (js/sim.js:36,44,738), set from
the iterations the last step ran as the frame's report reads them
(js/sim.js:1916-1920)."""
import ctypes as ct

import numpy as np

from . import gpu
from .gpu import WigError, _check

AIR_CYCLES = 64          # js/sim.js:36
AIR_CYCLES_SPARE = 8     # js/sim.js:44
AIR_TOLERANCE = 1e-4     # shaders/air.wgsl:945
PART_MG, PART_CG, PART_STATUS, PART_MEAN = 0, 1, 2, 3


class WigAirLayout(ct.Structure):
    _fields_ = [
        ("width", ct.c_uint32), ("height", ct.c_uint32), ("levels", ct.c_uint32), ("groups", ct.c_uint32),
        ("level_width", ct.c_uint32 * 9), ("level_height", ct.c_uint32 * 9), ("level_base", ct.c_uint32 * 9),
        ("part_bytes", ct.c_uint64 * 4),
    ]


_ready = False


def lib():
    global _ready
    L = gpu.lib()
    if not _ready:
        P = ct.c_void_p
        L.wig_air_create.restype = P
        L.wig_air_create.argtypes = [ct.c_uint32, ct.c_uint32]
        L.wig_air_destroy.argtypes = [P]
        L.wig_air_layout.argtypes = [P, ct.POINTER(WigAirLayout)]
        L.wig_air_upload.argtypes = [P, ct.c_int, P, ct.c_size_t]
        L.wig_air_download.argtypes = [P, ct.c_int, P, ct.c_size_t]
        L.wig_air_device_ptr.argtypes = [P, ct.c_int, ct.POINTER(P)]
        L.wig_air_set_stream.argtypes = [P, P]
        L.wig_air_set_tile.argtypes = [P, ct.c_uint32, ct.c_uint32]
        L.wig_air_set_coarse.argtypes = [P, ct.c_int]
        L.wig_air_project.argtypes = [P, ct.c_uint32]
        L.wig_air_sync.argtypes = [P]
        L.wig_air_time.argtypes = [P, ct.c_int]
        L.wig_air_elapsed.argtypes = [P, ct.POINTER(ct.c_double), ct.POINTER(ct.c_uint32)]
        L.wig_air_project_static.argtypes = [P, P, P, P, ct.c_uint32, ct.c_uint32, ct.c_uint32, P]
        _ready = True
    return L


_DTYPES = {PART_MG: np.float32, PART_CG: np.float32, PART_STATUS: np.uint32, PART_MEAN: np.float32}


class GpuAir:
    """An air grid's projection buffers on the GPU (wig_air_create)."""

    def __init__(self, width, height):
        L = lib()
        if not L.wig_device_ok():
            raise WigError("no HIP device visible")
        self._a = L.wig_air_create(width, height)
        if not self._a:
            raise WigError(f"wig_air_create failed: {L.wig_last_error().decode()}")
        self.layout = WigAirLayout()
        _check(L.wig_air_layout(self._a, ct.byref(self.layout)))
        self.width, self.height, self.levels = width, height, self.layout.levels
        self.dims = [(self.layout.level_width[l], self.layout.level_height[l]) for l in range(self.levels)]

    def close(self):
        if getattr(self, "_a", None):
            lib().wig_air_destroy(self._a)
            self._a = None

    __del__ = close

    def download(self, part):
        n = self.layout.part_bytes[part]
        out = np.empty(n // 4, _DTYPES[part])
        _check(lib().wig_air_download(self._a, part, out.ctypes.data, n))
        return out

    def upload(self, part, array):
        a = np.ascontiguousarray(array, _DTYPES[part])
        if a.nbytes != self.layout.part_bytes[part]:
            raise WigError(f"part {part}: {a.nbytes} bytes, the grid's are {self.layout.part_bytes[part]}")
        _check(lib().wig_air_upload(self._a, part, a.ctypes.data, a.nbytes))

    def load(self, mg0, x0, mean=None):
        """Level 0 (gather's and divergence's) and the warm start, the rest
        of the buffers none."""
        n = self.width * self.height
        mg = np.zeros(self.layout.part_bytes[PART_MG] // 4, np.float32)
        mg[:8 * n] = np.asarray(mg0, np.float32).reshape(-1)
        cg = np.zeros(self.layout.part_bytes[PART_CG] // 4, np.float32)
        cg[:4 * n].reshape(n, 4)[:, 0] = np.asarray(x0, np.float32).reshape(-1)
        self.upload(PART_MG, mg)
        self.upload(PART_CG, cg)
        self.upload(PART_STATUS, np.zeros(8, np.uint32))
        self.upload(PART_MEAN, np.zeros(n, np.float32) if mean is None else np.asarray(mean, np.float32).reshape(-1))

    def set_tile(self, side, level=None):
        """A level's LDS tile side (8, 16 or 32); every level's if level is None."""
        for l in range(self.levels) if level is None else [level]:
            _check(lib().wig_air_set_tile(self._a, l, side))

    def set_coarse(self, on):
        """The coarse levels in one kernel, or a launch each (the default)."""
        _check(lib().wig_air_set_coarse(self._a, 1 if on else 0))

    def set_stream(self, stream):
        _check(lib().wig_air_set_stream(self._a, stream))

    def project(self, cycles=AIR_CYCLES):
        _check(lib().wig_air_project(self._a, cycles))

    def device_ptr(self, part):
        p = ct.c_void_p()
        _check(lib().wig_air_device_ptr(self._a, part, ct.byref(p)))
        return p.value

    def sync(self):
        _check(lib().wig_air_sync(self._a))

    def time(self, on):
        _check(lib().wig_air_time(self._a, 1 if on else 0))

    def elapsed(self):
        ms, n = ct.c_double(), ct.c_uint32()
        _check(lib().wig_air_elapsed(self._a, ct.byref(ms), ct.byref(n)))
        return ms.value, n.value


# the most cells wig_air_project_static takes (WIG_AIR_STATIC_MAX_CELLS)
STATIC_MAX_CELLS = 262144


def project_static(mg, cg, status, mean, width, height, cycles=AIR_CYCLES, stream=None):
    """wig_air_project_static: the projection on the caller's device buffers
    (addresses of the four parts, laid out as a width x height grid's), with
    the working arrays in the library's static device array.  Allocates
    nothing; calls on one device must not overlap."""
    _check(lib().wig_air_project_static(mg, cg, status, mean, width, height, cycles, stream))


class AirProjection:
    """The air's projection a step at a time, as the reference's host encodes
    it: `cycles` iterations (airCycles), AIR_CYCLES at first and after a step
    that used all it was given, else twice the last step's plus
    AIR_CYCLES_SPARE (js/sim.js:1916-1920).  `grid` is a GpuAir (or the
    tests' oracle grid, which has the same methods)."""

    def __init__(self, grid):
        self.grid = grid
        self.cycles = AIR_CYCLES
        self.last_iterations = None

    def step(self):
        given = self.cycles
        self.grid.project(given)
        most = int(self.grid.download(PART_STATUS)[4]) if hasattr(self.grid, "download") else int(self.grid.status[4])
        self.last_iterations = most
        self.cycles = AIR_CYCLES if most >= given else min(2 * most + AIR_CYCLES_SPARE, AIR_CYCLES)
        return most
