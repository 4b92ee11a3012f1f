"""Shared helpers of the tests: worlds from scenes, stepping a pair of
worlds in lock step, and comparisons of their buffers."""
import numpy as np

from worldinagrain_amd import abi, scenes
from worldinagrain_amd import elements as E
from worldinagrain_amd.host import Simulation, SolverSettings


def make_oracle(width, height, cells, coarse=0):
    from oracle.oracle import OracleWorld
    w = OracleWorld(width, height, coarse_levels=coarse, kinds=E.kinds(), dot_mode=1)
    w.upload(abi.PART_CELLS, cells)
    return w


def make_gpu(width, height, cells, coarse=0):
    from worldinagrain_amd.gpu import GpuWorld
    w = GpuWorld(width, height, coarse_levels=coarse, kinds=E.kinds())
    w.upload(abi.PART_CELLS, cells)
    return w


def reference_sim(world, width, height, **kw):
    return Simulation(world, width, height, **kw)


def config_sim(world, width, height, tol=1e-6, **kw):
    return Simulation(world, width, height, dx=1.0, earth_g=float(np.float32(9.81 * (1 / 120) ** 2)), dt=1 / 120,
                      solver=SolverSettings.relative(tol), **kw)


def particle_rows(world, part):
    """The live particles of a buffer as rows of 11 words (motion, the state's
    word, temperature and id, affine), sorted: the multiset, whatever order
    the slots hold them in."""
    p = world.download(part) if hasattr(world, "download") else world.part(part)
    p = np.asarray(p).view(np.uint32)
    cap = world.layout.capacity
    n = min(int(p[0]), cap)
    items = p[4:4 + 12 * cap].reshape(3, cap, 4)
    rows = np.concatenate([items[0, :n], items[1, :n, :3], items[2, :n]], axis=1)
    order = np.lexsort(rows.T[::-1])
    return int(p[0]), rows[order]


def cell_parts(world):
    return {p: np.asarray(world.download(p)).view(np.uint32).copy()
            for p in (abi.PART_CELLS, abi.PART_COUNTERS, abi.PART_FIELD, abi.PART_BLOCKS, abi.PART_FLUIDCG)}


def first_difference(a, b, name):
    a = np.asarray(a).view(np.uint32)
    b = np.asarray(b).view(np.uint32)
    if a.shape != b.shape:
        return f"{name}: shapes {a.shape} vs {b.shape}"
    bad = np.nonzero(a != b)[0]
    if bad.size == 0:
        return None
    i = bad[0]
    return f"{name}: {bad.size} words differ, first at {i}: {a[i]:#x} vs {b[i]:#x}"


def compare_worlds(wa, wb, exact_particles=True):
    """Differences between two worlds' buffers (empty list: bit-identical,
    particles compared as multisets)."""
    diffs = []
    pa, pb = cell_parts(wa), cell_parts(wb)
    for p in pa:
        d = first_difference(pa[p], pb[p], abi.PART_NAMES[p])
        if d:
            diffs.append(d)
    for part in (abi.PART_PARTICLES0, abi.PART_PARTICLES1):
        na, ra = particle_rows(wa, part)
        nb, rb = particle_rows(wb, part)
        if na != nb:
            diffs.append(f"{abi.PART_NAMES[part]} count {na} vs {nb}")
        elif exact_particles:
            d = first_difference(ra.ravel(), rb.ravel(), abi.PART_NAMES[part] + " (as a multiset)")
            if d:
                diffs.append(d)
    return diffs
