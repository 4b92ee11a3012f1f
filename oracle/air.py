"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's
cpu_baseline leg).  ctypes binding of the air projection's CPU restatement
(oracle/air_oracle.c, in libwig_oracle.so): the reference's buffers (mg, cg,
status) as numpy arrays."""
import ctypes as ct

import numpy as np

from oracle.oracle import lib as _lib

_F = ct.POINTER(ct.c_float)
_U = ct.POINTER(ct.c_uint32)
_ready = False


def lib():
    global _ready
    L = _lib()
    if not _ready:
        L.wo_air_levels.restype = ct.c_uint32
        L.wo_air_levels.argtypes = [ct.c_uint32, ct.c_uint32, _U, _U]
        L.wo_air_mg_floats.restype = ct.c_uint64
        L.wo_air_mg_floats.argtypes = [ct.c_uint32, ct.c_uint32]
        L.wo_air_cg_floats.restype = ct.c_uint64
        L.wo_air_cg_floats.argtypes = [ct.c_uint32, ct.c_uint32]
        L.wo_air_clear_status.argtypes = [_U]
        L.wo_air_start.argtypes = [ct.c_uint32, ct.c_uint32, _F, _F, _U, _F]
        L.wo_air_iteration.argtypes = [ct.c_uint32, ct.c_uint32, _F, _F, _U, ct.c_uint32]
        L.wo_air_vcycle.argtypes = [ct.c_uint32, ct.c_uint32, _F, _F]
        L.wo_air_project.argtypes = [ct.c_uint32, ct.c_uint32, _F, _F, _U, _F, ct.c_uint32]
        _ready = True
    return L


def _p(a, t):
    return a.ctypes.data_as(t)


class OracleAir:
    """The air's projection buffers of a width x height air grid, stepped by
    the oracle."""

    def __init__(self, width, height):
        L = lib()
        self.width, self.height = width, height
        lw, lh = (ct.c_uint32 * 9)(), (ct.c_uint32 * 9)()
        self.levels = L.wo_air_levels(width, height, lw, lh)
        self.dims = [(lw[l], lh[l]) for l in range(self.levels)]
        self.mg = np.zeros(L.wo_air_mg_floats(width, height), np.float32)
        self.cg = np.zeros(L.wo_air_cg_floats(width, height), np.float32)
        self.status = np.zeros(8, np.uint32)
        self.mean = np.zeros(width * height, np.float32)

    def load(self, mg0, x0, mean=None):
        """Level 0 (gather's) and the warm start: the rest as allocated (none)."""
        n = self.width * self.height
        self.mg[:8 * n] = np.asarray(mg0, np.float32).reshape(-1)
        self.cg[:4 * n].reshape(n, 4)[:, 0] = np.asarray(x0, np.float32).reshape(-1)
        self.mean[:] = 0.0 if mean is None else np.asarray(mean, np.float32).reshape(-1)

    def project(self, cycles=64):
        lib().wo_air_project(self.width, self.height, _p(self.mg, _F), _p(self.cg, _F), _p(self.status, _U),
                             _p(self.mean, _F), cycles)

    def start(self):
        L = lib()
        L.wo_air_clear_status(_p(self.status, _U))
        L.wo_air_start(self.width, self.height, _p(self.mg, _F), _p(self.cg, _F), _p(self.status, _U),
                       _p(self.mean, _F))

    def iteration(self, cycle):
        lib().wo_air_iteration(self.width, self.height, _p(self.mg, _F), _p(self.cg, _F), _p(self.status, _U), cycle)

    def vcycle(self):
        lib().wo_air_vcycle(self.width, self.height, _p(self.mg, _F), _p(self.cg, _F))

    def level_base(self, l):
        """levelBase(l) in floats (shaders/air.wgsl:180-188)."""
        return 8 * sum(w * h for w, h in self.dims[:l])

    def level(self, l):
        """Level l's cells as a (h, w, 8) view."""
        w, h = self.dims[l]
        b = self.level_base(l)
        return self.mg[b:b + 8 * w * h].reshape(h, w, 8)

    def cells(self):
        """cg's per-cell words as a (h, w, 4) view (x, r, p, A p)."""
        n = self.width * self.height
        return self.cg[:4 * n].reshape(self.height, self.width, 4)

    def sums(self):
        return self.cg[4 * self.width * self.height:]
