"""ctypes binding of libwig.so (include/wig.h): a world whose buffers live on
the GPU and whose steps run the HIP kernels.  There is no CPU fallback: if
the library is missing or no device is visible, this raises."""
import ctypes as ct
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwig.so")
_lib = None


class WigError(RuntimeError):
    pass


def lib():
    """Loads libwig.so (built by worldinagrain_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise WigError(f"libwig.so is not built ({LIB_PATH}); run `python -m worldinagrain_amd.build`")
        L = ct.CDLL(LIB_PATH)
        L.wig_create.restype = ct.c_void_p
        L.wig_create.argtypes = [ct.c_uint32, ct.c_uint32, ct.c_uint32, ct.c_uint32]
        L.wig_destroy.argtypes = [ct.c_void_p]
        L.wig_set_kinds.argtypes = [ct.c_void_p, ct.c_char_p]
        L.wig_set_element_props.argtypes = [ct.c_void_p, ct.c_void_p]
        L.wig_layout.argtypes = [ct.c_void_p, ct.POINTER(abi.WigLayoutInfo)]
        L.wig_upload.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_size_t]
        L.wig_download.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_size_t]
        L.wig_device_ptr.argtypes = [ct.c_void_p, ct.c_int, ct.POINTER(ct.c_void_p)]
        L.wig_set_stream.argtypes = [ct.c_void_p, ct.c_void_p]
        L.wig_fluid_step.argtypes = [ct.c_void_p, ct.POINTER(abi.WigParams), ct.POINTER(abi.WigStepCtl)]
        L.wig_report.argtypes = [ct.c_void_p, ct.POINTER(ct.c_uint32)]
        L.wig_report_async.argtypes = [ct.c_void_p, ct.POINTER(ct.c_uint64)]
        L.wig_report_take.argtypes = [ct.c_void_p, ct.c_uint64, ct.POINTER(ct.c_uint32), ct.c_int]
        L.wig_solve_stats.argtypes = [ct.c_void_p, ct.POINTER(abi.WigSolveStats)]
        L.wig_sync.argtypes = [ct.c_void_p]
        L.wig_last_error.restype = ct.c_char_p
        L.wig_device_ok.restype = ct.c_int
        L.wig_run_totals.argtypes = [ct.c_void_p, ct.POINTER(abi.WigRunTotals)]
        L.wig_time_kernel.argtypes = [ct.c_void_p, ct.c_int]
        L.wig_kernel_time.argtypes = [ct.c_void_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_uint32)]
        L.wig_kernel_name.restype = ct.c_char_p
        L.wig_kernel_name.argtypes = [ct.c_int]
        L.wig_kernel_count.restype = ct.c_int
        L.wig_set_tile.argtypes = [ct.c_void_p, ct.c_uint32, ct.c_uint32]
        L.wig_set_lds_level.argtypes = [ct.c_void_p, ct.c_uint32]
        L.wig_set_coarse_wave.argtypes = [ct.c_void_p, ct.c_uint32]
        L.wig_batch_create.restype = ct.c_void_p
        L.wig_batch_create.argtypes = [ct.POINTER(ct.c_void_p), ct.c_uint32]
        L.wig_batch_destroy.argtypes = [ct.c_void_p]
        L.wig_batch_assign.argtypes = [ct.c_void_p, ct.POINTER(ct.c_void_p), ct.c_uint32]
        L.wig_batch_set_stream.argtypes = [ct.c_void_p, ct.c_void_p]
        L.wig_batch_step.argtypes = [ct.c_void_p, ct.POINTER(abi.WigParams), ct.POINTER(abi.WigStepCtl)]
        L.wig_batch_report.argtypes = [ct.c_void_p, ct.POINTER(ct.c_uint32)]
        L.wig_batch_report_async.argtypes = [ct.c_void_p, ct.POINTER(ct.c_uint64)]
        L.wig_batch_report_take.argtypes = [ct.c_void_p, ct.c_uint64, ct.POINTER(ct.c_uint32), ct.c_int]
        L.wig_batch_sync.argtypes = [ct.c_void_p]
        L.wig_batch_time_kernel.argtypes = [ct.c_void_p, ct.c_int]
        L.wig_set_graphs.argtypes = [ct.c_void_p, ct.c_int]
        L.wig_capture.argtypes = [ct.c_void_p, ct.c_int]
        L.wig_cg64.argtypes = [ct.c_void_p, ct.c_double, ct.c_uint32, ct.POINTER(abi.WigCheck64)]
        L.wig_cg64_x.argtypes = [ct.c_void_p, ct.POINTER(ct.c_double)]
        L.wig_batch_set_graphs.argtypes = [ct.c_void_p, ct.c_int]
        L.wig_batch_kernel_time.argtypes = [ct.c_void_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_uint32)]
        _lib = L
    return _lib


def _check(status):
    if status < 0:
        raise WigError(f"wig error {status}: {lib().wig_last_error().decode()}")
    return status


_DTYPES = {abi.PART_FIELD: np.float32, abi.PART_FLUIDCG: np.float32, abi.PART_SOLVE: np.uint8, abi.PART_AIR: np.float32,
           abi.PART_GRAVITY: np.float32}


class HipStream:
    """A HIP stream (default priority) of the HIP runtime libwig.so runs on,
    for set_stream and the serving loop's copy stream.  (A process holds one
    runtime if torch, which carries its own, is imported before libwig.so
    loads: then torch's streams serve as well.  Loaded the other way round,
    the two runtimes' streams do not mix, and these are the ones to use.)"""

    def __init__(self):
        lib()
        self._hip = ct.CDLL("libamdhip64.so.7")         # the runtime libwig.so loaded
        self._hip.hipStreamCreate.argtypes = [ct.POINTER(ct.c_void_p)]
        self._hip.hipStreamDestroy.argtypes = [ct.c_void_p]
        h = ct.c_void_p()
        if self._hip.hipStreamCreate(ct.byref(h)) != 0:
            raise WigError("hipStreamCreate failed")
        self.handle = h.value

    def close(self):
        if self.handle:
            self._hip.hipStreamDestroy(ct.c_void_p(self.handle))
            self.handle = None


class GpuWorld:
    """A world on the GPU (wig_create); the same interface as the tests' oracle world."""

    def __init__(self, width, height, coarse_levels=0, kinds=None):
        L = lib()
        if not L.wig_device_ok():
            raise WigError("no HIP device visible")
        self._w = L.wig_create(width, height, coarse_levels, 0)
        if not self._w:
            raise WigError(f"wig_create failed: {L.wig_last_error().decode()}")
        self.layout = abi.WigLayoutInfo()
        _check(L.wig_layout(self._w, ct.byref(self.layout)))
        self.kinds = None
        self.knobs = {}         # speed settings made (tiles, coarse levels); a clone takes them
        if kinds is not None:
            self.set_kinds(kinds)

    def close(self):
        if getattr(self, "_w", None):
            lib().wig_destroy(self._w)
            self._w = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_kinds(self, kinds):
        k = np.ascontiguousarray(kinds, dtype=np.uint8)
        assert k.shape == (256,)
        _check(lib().wig_set_kinds(self._w, k.tobytes()))
        self.kinds = k

    def set_element_props(self, props):
        """The elements' columns (wig_set_element_props): (256, 6) f32,
        viscosity, viscosity_e, viscosity_t, tension, wetting, breaks (fewer
        columns padded with 0)."""
        p = np.ascontiguousarray(abi.element_props(props))
        _check(lib().wig_set_element_props(self._w, p.ctypes.data))

    def part_bytes(self, p):
        return int(self.layout.part_bytes[p])

    def upload(self, p, data):
        buf = np.ascontiguousarray(data)
        assert buf.nbytes == self.part_bytes(p), (p, buf.nbytes, self.part_bytes(p))
        _check(lib().wig_upload(self._w, p, buf.ctypes.data, buf.nbytes))

    def download(self, p):
        dt = _DTYPES.get(p, np.uint32)
        out = np.empty(self.part_bytes(p) // np.dtype(dt).itemsize, dtype=dt)
        _check(lib().wig_download(self._w, p, out.ctypes.data, out.nbytes))
        return out

    part = download

    def device_ptr(self, p):
        ptr = ct.c_void_p()
        _check(lib().wig_device_ptr(self._w, p, ct.byref(ptr)))
        return ptr.value

    def set_stream(self, stream_ptr):
        _check(lib().wig_set_stream(self._w, ct.c_void_p(stream_ptr) if stream_ptr else None))

    def step(self, params, ctl):
        return _check(lib().wig_fluid_step(self._w, ct.byref(params), ct.byref(ctl)))

    def report(self):
        stats = (ct.c_uint32 * abi.STATS)()
        _check(lib().wig_report(self._w, stats))
        return np.array(stats, dtype=np.uint32)

    def report_async(self):
        """The report asked for without waiting (wig_report_async): its ticket."""
        t = ct.c_uint64()
        _check(lib().wig_report_async(self._w, ct.byref(t)))
        return t.value

    def report_take(self, ticket, wait=True):
        """The report of `ticket` (wig_report_take); None if it has not landed and not `wait`."""
        stats = (ct.c_uint32 * abi.STATS)()
        st = lib().wig_report_take(self._w, ct.c_uint64(ticket), stats, 1 if wait else 0)
        if st == abi.WIG_NOT_READY:
            return None
        _check(st)
        return np.array(stats, dtype=np.uint32)

    def solve_stats(self):
        s = abi.WigSolveStats()
        _check(lib().wig_solve_stats(self._w, ct.byref(s)))
        return s

    def sync(self):
        _check(lib().wig_sync(self._w))

    def run_totals(self):
        t = abi.WigRunTotals()
        _check(lib().wig_run_totals(self._w, ct.byref(t)))
        return t

    def time_kernel(self, name):
        """Brackets every later launch of kernel `name` with HIP events (None stops)."""
        if name is None:
            _check(lib().wig_time_kernel(self._w, -1))
            return
        _check(lib().wig_time_kernel(self._w, kernel_id(name)))

    def set_tile(self, level, side):
        """The coarse-level sweep kernels' tile side on `level` (wig_set_tile)."""
        _check(lib().wig_set_tile(self._w, level, side))
        self.knobs[("tile", level)] = side

    def set_graphs(self, on):
        """Steps replayed as hipGraphs (wig_set_graphs)."""
        _check(lib().wig_set_graphs(self._w, 1 if on else 0))

    def set_lds_level(self, level):
        """Where the one-workgroup coarse-level kernel starts (wig_set_lds_level)."""
        _check(lib().wig_set_lds_level(self._w, level))
        self.knobs[("lds_level",)] = level

    def set_coarse_wave(self, cells):
        """The most cells a coarse level below the one-workgroup kernel's
        first may have to be worked by one wavefront (wig_set_coarse_wave)."""
        _check(lib().wig_set_coarse_wave(self._w, cells))
        self.knobs[("coarse_wave",)] = cells

    def capture(self, on=True):
        """Each later solve keeps its right-hand side, for cg64 (wig_capture)."""
        _check(lib().wig_capture(self._w, 1 if on else 0))

    def cg64(self, tol=1e-12, maxit=200000):
        """FP64 CG on the last solve's matrix and rhs, and the solve's pressure
        measured against it (wig_cg64): (WigCheck64, x64 per cell)."""
        out = abi.WigCheck64()
        _check(lib().wig_cg64(self._w, tol, maxit, ct.byref(out)))
        x = np.empty(self.layout.cells, dtype=np.float64)
        _check(lib().wig_cg64_x(self._w, x.ctypes.data_as(ct.POINTER(ct.c_double))))
        return out, x

    def clone(self):
        """A new world of the same size and kinds holding this one's parts,
        with its speed settings."""
        w = GpuWorld(self.layout.width, self.layout.height, coarse_levels=self.layout.levels, kinds=self.kinds)
        for p in range(abi.PART_COUNT):
            w.upload(p, self.download(p))
        lds = self.knobs.get(("lds_level",))
        if lds is not None:
            w.set_lds_level(lds)        # (before the tiles: it decides which levels are tiled)
        for key, v in self.knobs.items():
            if key[0] == "tile":
                try:
                    w.set_tile(key[1], v)
                except WigError:
                    pass        # (a level the one-workgroup kernel has since taken)
            elif key[0] == "coarse_wave":
                w.set_coarse_wave(v)
        return w

    def kernel_time(self):
        """(ms summed, launches) of the timed kernel since time_kernel."""
        ms, n = ct.c_double(), ct.c_uint32()
        _check(lib().wig_kernel_time(self._w, ct.byref(ms), ct.byref(n)))
        return ms.value, n.value


class GpuBatch:
    """Worlds stepped together, each kernel of a step one launch over all of
    them (wig_batch_*).  The worlds stay the caller's."""

    def __init__(self, worlds):
        L = lib()
        self.worlds = list(worlds)
        arr = (ct.c_void_p * len(self.worlds))(*[w._w for w in self.worlds])
        self._b = L.wig_batch_create(arr, len(self.worlds))
        if not self._b:
            raise WigError(f"wig_batch_create failed: {L.wig_last_error().decode()}")

    def close(self):
        if getattr(self, "_b", None):
            lib().wig_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.worlds)

    def assign(self, worlds):
        """The batch's worlds replaced (wig_batch_assign)."""
        self.worlds = list(worlds)
        arr = (ct.c_void_p * len(self.worlds))(*[w._w for w in self.worlds])
        _check(lib().wig_batch_assign(self._b, arr, len(self.worlds)))

    def set_stream(self, stream_ptr):
        _check(lib().wig_batch_set_stream(self._b, ct.c_void_p(stream_ptr) if stream_ptr else None))

    def set_graphs(self, on):
        _check(lib().wig_batch_set_graphs(self._b, 1 if on else 0))

    def step(self, params, ctls):
        n = len(self.worlds)
        assert len(params) == n and len(ctls) == n
        P = (abi.WigParams * n)(*params)
        C = (abi.WigStepCtl * n)(*ctls)
        return _check(lib().wig_batch_step(self._b, P, C))

    def step_arrays(self, P, C):
        """wig_batch_step over ready ctypes arrays of WigParams and WigStepCtl."""
        return _check(lib().wig_batch_step(self._b, P, C))

    def report(self):
        n = len(self.worlds)
        out = (ct.c_uint32 * (abi.STATS * n))()
        _check(lib().wig_batch_report(self._b, out))
        return np.array(out, dtype=np.uint32).reshape(n, abi.STATS)

    def report_async(self):
        t = ct.c_uint64()
        _check(lib().wig_batch_report_async(self._b, ct.byref(t)))
        return t.value

    def report_take(self, ticket, wait=True):
        n = len(self.worlds)
        out = (ct.c_uint32 * (abi.STATS * n))()
        st = lib().wig_batch_report_take(self._b, ct.c_uint64(ticket), out, 1 if wait else 0)
        if st == abi.WIG_NOT_READY:
            return None
        _check(st)
        return np.array(out, dtype=np.uint32).reshape(n, abi.STATS)

    def sync(self):
        _check(lib().wig_batch_sync(self._b))

    def time_kernel(self, name):
        _check(lib().wig_batch_time_kernel(self._b, -1 if name is None else kernel_id(name)))

    def kernel_time(self):
        ms, n = ct.c_double(), ct.c_uint32()
        _check(lib().wig_batch_kernel_time(self._b, ct.byref(ms), ct.byref(n)))
        return ms.value, n.value


def kernel_names():
    L = lib()
    return [L.wig_kernel_name(i).decode() for i in range(L.wig_kernel_count())]


def kernel_id(name):
    return kernel_names().index(name)
