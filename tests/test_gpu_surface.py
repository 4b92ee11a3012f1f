"""Surface tension in the liquid step on the GPU against the oracle: a
world of millimetre cells (the capillary length about three cells), a drop
falling into a pool in a glass the water wets, so the surface pass sets σκ
on the open cells beside the liquid, classify damps the surface with its
viscosity, and the sweeps run every step (js/sim.js:2449-2514).  Every
buffer bit for bit; then the same world in a batch beside a centimetre
world without tension, each as it steps alone."""
import sys

import numpy as np
import pytest

from worldinagrain_amd import abi, scenes, elements as E
from worldinagrain_amd.host import Simulation, SolverSettings, step_batch, surface_tension
from tests.helpers import compare_worlds

pytestmark = pytest.mark.gpu

W, H = 80, 64
DX = 0.001                      # a millimetre a cell


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def props():
    p = np.zeros((256, 5), dtype=np.float32)
    p[E.WATER, 3] = 1.0          # water's tension
    p[E.WALL, 4] = 0.8           # the glass: wetted, cos θ 0.6
    return p


def scene():
    cells = scenes.empty_cells(W, H)
    yy, xx = np.mgrid[0:H, 0:W]
    glass = ((xx == 10) | (xx == 69) | (yy == 60)) & (xx >= 10) & (xx <= 69) & (yy >= 20) & (yy <= 60)
    scenes.paint(cells, W, H, glass | scenes.border(W, H), E.WALL)
    pool = (xx > 10) & (xx < 69) & (yy > 44) & (yy < 60)
    drop = (xx - 40) ** 2 + (yy - 28) ** 2 <= 25
    scenes.paint(cells, W, H, pool | drop, E.WATER)
    return cells


def sim_of(w, dx=DX):
    g = float(np.float32(9.81 * (1 / 240) ** 2 / dx))
    return Simulation(w, W, H, dx=dx, earth_g=g, dt=1 / 240, solver=SolverSettings.relative(1e-6))


def world(kind, cells):
    if kind == "oracle":
        from oracle.oracle import OracleWorld
        w = OracleWorld(W, H, kinds=E.kinds(), dot_mode=1)
    else:
        from worldinagrain_amd.gpu import GpuWorld
        w = GpuWorld(W, H, kinds=E.kinds())
    w.set_element_props(props())
    w.upload(abi.PART_CELLS, cells)
    return w


def test_surface_tension_matches_oracle():
    cells = scene()
    wo, wg = world("oracle", cells), world("gpu", cells)
    so, sg = sim_of(wo), sim_of(wg)
    assert surface_tension(DX) and not surface_tension(0.01)
    for s in range(12):
        so.step()
        sg.step()
        assert np.array_equal(so.last_stats, sg.last_stats), (s, so.last_stats, sg.last_stats)
        d = compare_worlds(wo, wg)
        assert not d, f"step {s}: " + "; ".join(d)
    # open cells beside the liquid hold σκ
    f = wg.download(abi.PART_FIELD)[:8 * W * H].reshape(H, W, 8)
    open_ = f[..., 2] == 0.0
    assert np.any(f[..., 0][open_] != 0.0)
    say(f"[surface] 12 steps, GPU == oracle; {int(sg.last_stats[abi.STAT_PARTICLES])} particles")


def test_surface_batch_matches_alone():
    from worldinagrain_amd.gpu import GpuBatch
    cells = scene()
    pairs = []
    for dx in (DX, 0.01):
        wb, wa = world("gpu", cells), world("gpu", cells)
        pairs.append((wb, sim_of(wb, dx), wa, sim_of(wa, dx)))
    batch = GpuBatch([p[0] for p in pairs])
    for s in range(5):
        step_batch(batch, [p[1] for p in pairs])
        for wb, sb, wa, sa in pairs:
            sa.step()
            assert np.array_equal(sb.last_stats, sa.last_stats), s
            d = compare_worlds(wb, wa)
            assert not d, f"step {s}: " + "; ".join(d)
