"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's
cpu_baseline leg).  ctypes binding of libwig_oracle.so, the CPU restatement of
the reference's liquid step (oracle/wig_oracle.c)."""
import ctypes as ct
import os

import numpy as np

from worldinagrain_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "libwig_oracle.so")
_libs = {}

# one pass of the step (wor_run's WOR_PASS_* ids, wig_oracle.c)
(PASS_SYNC, PASS_SYNC_PARTICLES, PASS_SYNC_SEED, PASS_COMPACT, PASS_RETIRE, PASS_CLEAR, PASS_GATHER, PASS_SEAL,
 PASS_CLASSIFY, PASS_SET_FACES, PASS_UNPRESS, PASS_DIVERGENCE, PASS_RESTRICT, PASS_SMOOTH, PASS_COARSEST, PASS_PROLONG,
 PASS_PROLONG_FINE, PASS_PROJECT, PASS_EXTRAPOLATE, PASS_ADVECT, PASS_SORT, PASS_RASTER_CLAIMS, PASS_RASTER,
 PASS_UNPAINT, PASS_RELAX, PASS_VISCOSITY, PASS_SURFACE) = range(27)


OMP_PATH = os.path.join(_HERE, "libwig_oracle_omp.so")


def mutant_path(k):
    """The oracle built with planted mistake k (`make -C oracle mutants`)."""
    return os.path.join(_HERE, "_mut", f"libwig_oracle_m{k}.so")


def lib(path=None):
    """The oracle's library (or, given a path, that build of it)."""
    path = path or _PATH
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = ct.CDLL(path)
        L.wor_create.restype = ct.c_void_p
        L.wor_create.argtypes = [ct.c_uint32, ct.c_uint32, ct.c_uint32]
        L.wor_destroy.argtypes = [ct.c_void_p]
        L.wor_part.restype = ct.c_void_p
        L.wor_part.argtypes = [ct.c_void_p, ct.c_int, ct.POINTER(ct.c_uint64)]
        L.wor_set_kinds.argtypes = [ct.c_void_p, ct.c_char_p]
        L.wor_set_element_props.argtypes = [ct.c_void_p, ct.c_void_p]
        L.wor_set_dot_mode.argtypes = [ct.c_void_p, ct.c_int]
        L.wor_layout.argtypes = [ct.c_void_p, ct.POINTER(abi.WigLayoutInfo)]
        L.wor_fluid_step.argtypes = [ct.c_void_p, ct.POINTER(abi.WigParams), ct.POINTER(abi.WigStepCtl)]
        L.wor_capture.argtypes = [ct.c_void_p, ct.c_int]
        L.wor_captured_b.restype = ct.POINTER(ct.c_float)
        L.wor_captured_b.argtypes = [ct.c_void_p]
        L.wor_cg64.argtypes = [ct.c_void_p, ct.POINTER(ct.c_float), ct.POINTER(ct.c_double), ct.c_double, ct.c_int,
                               ct.POINTER(ct.c_double)]
        L.wor_relres.restype = ct.c_double
        L.wor_relres.argtypes = [ct.c_void_p, ct.POINTER(ct.c_float), ct.POINTER(ct.c_float)]
        L.wor_pcg.restype = ct.c_uint32
        L.wor_pcg.argtypes = [ct.c_uint32]
        L.wor_seed_id.restype = ct.c_uint32
        L.wor_seed_id.argtypes = [ct.c_void_p, ct.c_int32, ct.c_int32, ct.c_uint32, ct.c_uint32]
        L.wor_seed_pos.argtypes = [ct.c_int32, ct.c_int32, ct.c_uint32, ct.c_uint32, ct.c_uint32, ct.c_uint32,
                                   ct.POINTER(ct.c_float)]
        L.wor_f16.restype = ct.c_uint16
        L.wor_f16.argtypes = [ct.c_float]
        L.wor_f32.restype = ct.c_float
        L.wor_f32.argtypes = [ct.c_uint16]
        L.wor_exact_parts.argtypes = [ct.c_float, ct.c_float, ct.POINTER(ct.c_uint32)]
        L.wor_exact_value.restype = ct.c_float
        L.wor_exact_value.argtypes = [ct.c_uint32, ct.c_uint32, ct.c_float]
        L.wor_fine_beta.restype = ct.c_float
        L.wor_fine_beta.argtypes = [ct.c_float] * 4
        L.wor_run.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.POINTER(abi.WigParams), ct.POINTER(abi.WigStepCtl)]
        L.wor_body_probe.argtypes = [ct.c_void_p, ct.POINTER(abi.WigParams), ct.POINTER(abi.WigStepCtl), ct.c_int,
                                     ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.wor_level_start.restype = ct.c_uint64
        L.wor_level_start.argtypes = [ct.c_void_p, ct.c_uint32]
        _libs[path] = L
    return _libs[path]


_DTYPES = {abi.PART_FIELD: np.float32, abi.PART_FLUIDCG: np.float32, abi.PART_SOLVE: np.uint8, abi.PART_AIR: np.float32,
           abi.PART_GRAVITY: np.float32}


class OracleWorld:
    """A world in host memory, stepped by the CPU restatement."""

    def __init__(self, width, height, coarse_levels=0, kinds=None, dot_mode=1, L=None):
        L = L or lib()
        self.L = L
        self._w = L.wor_create(width, height, coarse_levels)
        if not self._w:
            raise ValueError("invalid world size")
        self.layout = abi.WigLayoutInfo()
        L.wor_layout(self._w, ct.byref(self.layout))
        if kinds is not None:
            self.set_kinds(kinds)
        L.wor_set_dot_mode(self._w, dot_mode)

    def close(self):
        if self._w:
            self.L.wor_destroy(self._w)
            self._w = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_kinds(self, kinds):
        k = np.asarray(kinds, dtype=np.uint8)
        assert k.shape == (256,)
        self.L.wor_set_kinds(self._w, k.tobytes())

    def set_element_props(self, props):
        """(256, 6) f32: viscosity, viscosity_e, viscosity_t, tension, wetting,
        breaks (fewer columns padded with 0)."""
        p = np.ascontiguousarray(abi.element_props(props))
        self.L.wor_set_element_props(self._w, p.ctypes.data)

    def part(self, p):
        """A numpy view (u32 words, f32 for field/fluidCg) of the part's memory."""
        n = ct.c_uint64()
        ptr = self.L.wor_part(self._w, p, ct.byref(n))
        dt = _DTYPES.get(p, np.uint32)
        count = n.value // np.dtype(dt).itemsize
        buf = (ct.c_char * n.value).from_address(ptr)
        return np.frombuffer(buf, dtype=dt, count=count)

    def upload(self, p, data):
        dst = self.part(p)
        src = np.frombuffer(np.ascontiguousarray(data).tobytes(), dtype=dst.dtype)
        assert src.size == dst.size, (p, src.size, dst.size)
        dst[:] = src

    def download(self, p):
        return self.part(p).copy()

    def step(self, params, ctl):
        return self.L.wor_fluid_step(self._w, ct.byref(params), ct.byref(ctl))

    def capture(self, on=True):
        self.L.wor_capture(self._w, 1 if on else 0)

    def captured_b(self):
        ptr = self.L.wor_captured_b(self._w)
        return np.ctypeslib.as_array(ptr, shape=(self.layout.cells,)).copy()

    def cg64(self, b, tol=1e-12, maxit=100000):
        b = np.ascontiguousarray(b, dtype=np.float32)
        x = np.zeros(self.layout.cells, dtype=np.float64)
        rel = ct.c_double()
        it = self.L.wor_cg64(self._w, b.ctypes.data_as(ct.POINTER(ct.c_float)),
                            x.ctypes.data_as(ct.POINTER(ct.c_double)), tol, maxit, ct.byref(rel))
        return x, it, rel.value

    def relres(self, b, x):
        b = np.ascontiguousarray(b, dtype=np.float32)
        x = np.ascontiguousarray(x, dtype=np.float32)
        return self.L.wor_relres(self._w, b.ctypes.data_as(ct.POINTER(ct.c_float)), x.ctypes.data_as(ct.POINTER(ct.c_float)))

    def run(self, pass_id, arg, params, ctl):
        """One pass of the step on the buffers as they stand (wor_run)."""
        st = self.L.wor_run(self._w, pass_id, arg, ct.byref(params), ct.byref(ctl))
        assert st == 0, f"wor_run({pass_id}, {arg}): {st}"

    def body_probe(self, op, ints=(), floats=(), params=None, ctl=None):
        """One of the bodies' helpers (wor_body_probe's op): (floats out, u32 out)."""
        ia = (ct.c_int32 * 8)(*[int(v) for v in ints])
        fa = (ct.c_float * 8)(*[float(v) for v in floats])
        fo = (ct.c_float * 4)()
        uo = (ct.c_uint32 * 4)()
        params = params if params is not None else abi.WigParams()
        ctl = ctl if ctl is not None else abi.WigStepCtl()
        st = self.L.wor_body_probe(self._w, ct.byref(params), ct.byref(ctl), op, ia, fa, fo, uo)
        assert st == 0, f"wor_body_probe({op}): {st}"
        return list(fo), list(uo)

    def level_start(self, l):
        """Where coarse level l starts in the field, in vec4s."""
        return int(self.L.wor_level_start(self._w, l))

    def solve_stats(self):
        raw = self.part(abi.PART_SOLVE)[:64].tobytes()
        return abi.WigSolveStats.from_buffer_copy(raw)

    def report(self):
        """FLUID_STATS, then MOST_CYCLES cleared (js/sim.js:1760-1780)."""
        B = self.layout.blocks
        blocks = self.part(abi.PART_BLOCKS)
        stats = blocks[4 * B:4 * B + abi.STATS].copy()
        blocks[4 * B + abi.STAT_MOST_CYCLES] = 0
        return stats
