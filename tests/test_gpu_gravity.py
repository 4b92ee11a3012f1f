"""The gravity pass's field in the liquid step (WigStepCtl.gravity) on the
GPU against the oracle: a mass beside a dam break pulls the water sideways
and turns down near it (pull.wgsl's pullAt and downAt: the faces' pull, the
slip, the falling test, the bubble test, the escape and where displaced gas
goes), every buffer bit for bit; then the same world in a batch beside one
under Earth's pull alone, each as it steps alone."""
import sys

import numpy as np
import pytest

from worldinagrain_amd import abi, scenes
from worldinagrain_amd.host import step_batch
from tests.helpers import make_oracle, make_gpu, compare_worlds, config_sim

pytestmark = pytest.mark.gpu

W, H = 96, 80
G_STEP = float(np.float32(9.81 * (1 / 120) ** 2))


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def mass_field(mx=70.0, my=40.0, strength=0.4):
    """The field a mass at (mx, my) would give: its pull, 1/r^2 toward it,
    softened; its potential; down, the step nearest the total pull
    (common.wgsl's stepIndex order), DOWN_WEIGHTLESS where that is weak."""
    gw, gh = (W + 1) // 2, (H + 1) // 2
    y, x = np.mgrid[0:gh, 0:gw].astype(np.float64)
    px, py = 2.0 * x + 1.0, 2.0 * y + 1.0
    dx, dy = mx - px, my - py
    r2 = dx * dx + dy * dy + 16.0
    k = strength * G_STEP * 400.0 / r2
    fx, fy = k * dx / np.sqrt(r2), k * dy / np.sqrt(r2)
    tx, ty = fx, fy + G_STEP
    sector = np.rint(np.arctan2(ty, tx) / (np.pi / 4)).astype(np.int64) % 8
    down = sector + np.where(np.hypot(tx, ty) < 0.1 * G_STEP, abi.DOWN_WEIGHTLESS, 0)
    f = np.zeros((gh, gw, 4), dtype=np.float32)
    f[..., 0], f[..., 1] = fx, fy
    f[..., 2] = -k * np.sqrt(r2)
    f[..., 3] = down
    return f.reshape(-1)


def pulled(world, field):
    world.upload(abi.PART_GRAVITY, field)
    s = config_sim(world, W, H, tol=1e-6)
    s.gravity = True
    return s


def test_gravity_field_matches_oracle():
    cells = scenes.dam_break(W, H, block=(40, 60))
    field = mass_field()
    assert len(np.unique(field.reshape(-1, 4)[:, 3] % 8)) >= 3      # down turns across the world
    wo, wg = make_oracle(W, H, cells), make_gpu(W, H, cells)
    so, sg = pulled(wo, field), pulled(wg, field)
    for s in range(12):
        so.step()
        sg.step()
        assert np.array_equal(so.last_stats, sg.last_stats), (s, so.last_stats, sg.last_stats)
        d = compare_worlds(wo, wg)
        assert not d, f"step {s}: " + "; ".join(d)
    # the field did something: the world differs from one under Earth's pull alone
    we = make_gpu(W, H, cells)
    se = config_sim(we, W, H, tol=1e-6)
    for _ in range(12):
        se.step()
    assert compare_worlds(wg, we), "the field changed nothing"
    say(f"[gravity] 12 steps, GPU == oracle; {int(sg.last_stats[abi.STAT_PARTICLES])} particles")


def test_gravity_batch_matches_alone():
    from worldinagrain_amd.gpu import GpuBatch
    cells = scenes.dam_break(W, H, block=(40, 60))
    field = mass_field()
    pairs = []
    for with_field in (True, False):
        wb, wa = make_gpu(W, H, cells), make_gpu(W, H, cells)
        if with_field:
            sb, sa = pulled(wb, field), pulled(wa, field)
        else:
            sb, sa = config_sim(wb, W, H, tol=1e-6), config_sim(wa, W, H, tol=1e-6)
        pairs.append((wb, sb, wa, sa))
    batch = GpuBatch([p[0] for p in pairs])
    for s in range(5):
        step_batch(batch, [p[1] for p in pairs])
        for wb, sb, wa, sa in pairs:
            sa.step()
            assert np.array_equal(sb.last_stats, sa.last_stats), s
            d = compare_worlds(wb, wa)
            assert not d, f"step {s}: " + "; ".join(d)
