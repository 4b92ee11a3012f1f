"""The liquid step with the air coupled in (WigStepCtl.air) on the GPU
against the oracle: gather's push of the air's pressure on particles beside
gas, setFaces' drag of the wind along the surface, and raster's displaced gas,
all read from the world's air[0] part and added into its airSource part as
exact sums.  Every buffer, airSource included, bit for bit, over a dam break
and a splash of discs thrown through humid gas; then the same worlds in one
batch beside a world without the air, each as it steps alone."""
import sys

import numpy as np
import pytest

from worldinagrain_amd import abi, scenes, elements as E
from worldinagrain_amd.host import step_batch
from tests.helpers import make_oracle, make_gpu, compare_worlds, config_sim

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def thrown(cells, W, H, vx, vy):
    """The water of `cells` given the velocity (vx, vy), so that it crosses
    into the gas within the first steps."""
    c = cells.copy()
    w = c[:4 * W * H].reshape(-1, 4)
    water = (w[:, 0] & 0xFF) == E.WATER
    packed = int(np.array([vx, vy], np.float16).view(np.uint32)[0])
    w[water, 2] = packed
    # humid gas: vapour, mist and some oxygen used up beside its amount
    gas = c[4 * W * H:].reshape(-1, 2)[:W * H]
    held = gas[:, 1] != 0
    gas[held, 0] = int(np.array([0.0125, 0.003], np.float16).view(np.uint32)[0])
    gas[held, 1] = int(np.array([0.05, 1.0], np.float16).view(np.uint32)[0])
    return c


def spec():
    discs, _ = scenes.discs_dam(160, 120, discs=12, rmin=3, rmax=9)
    return [
        ("thrown dam 96x96", 96, 96, thrown(scenes.dam_break(96, 96, block=(48, 70)), 96, 96, 0.6, 0.25)),
        ("thrown discs 160x120", 160, 120, thrown(discs, 160, 120, -0.4, 0.5)),
    ]


def coupled(world, W, H, cells, seed):
    world.upload(abi.PART_AIR, scenes.air_over(W, H, cells, seed=seed, wind=0.2, lapse=0.01))
    world.upload(abi.PART_AIR_SOURCE, np.zeros(W * H * abi.SOURCE_WORDS, dtype=np.uint32))
    s = config_sim(world, W, H, tol=1e-6)
    s.air = True
    return s


def same_source(wa, wb, what):
    a = np.asarray(wa.download(abi.PART_AIR_SOURCE)).view(np.uint32)
    b = np.asarray(wb.download(abi.PART_AIR_SOURCE)).view(np.uint32)
    bad = np.nonzero(a != b)[0]
    assert bad.size == 0, f"{what}: airSource differs at {bad.size} words, first {bad[:6]}"
    return a


@pytest.mark.parametrize("which", [0, 1])
def test_air_coupled_step_matches_oracle(which):
    name, W, H, cells = spec()[which]
    say(f"[air coupling] {name}")
    wo, wg = make_oracle(W, H, cells), make_gpu(W, H, cells)
    so, sg = coupled(wo, W, H, cells, 7 + which), coupled(wg, W, H, cells, 7 + which)
    for s in range(8):
        so.step()
        sg.step()
        assert np.array_equal(so.last_stats, sg.last_stats), (s, so.last_stats, sg.last_stats)
        d = compare_worlds(wo, wg)
        assert not d, f"{name} step {s}: " + "; ".join(d)
        src = same_source(wo, wg, f"{name} step {s}")
    # the coupling did something: the air took drag on its faces and gas
    # the liquid displaced, and no particle was lost on the way
    v = src.reshape(H, W, abi.SOURCE_WORDS // 2, 2)
    moved = (v[..., abi.SOURCE_RIGHT, :] != 0).any() or (v[..., abi.SOURCE_DOWN, :] != 0).any()
    assert moved, "no drag reached the air"
    for k in (abi.SOURCE_AMOUNT, abi.SOURCE_VAPOUR, abi.SOURCE_MIST, abi.SOURCE_LOST, abi.SOURCE_HEAT):
        assert (v[..., k, :] != 0).any(), f"no gas was displaced (airSource slot {k})"
    say(f"[air coupling] {name}: 8 steps, GPU == oracle incl. airSource, "
        f"{int(sg.last_stats[abi.STAT_PARTICLES])} particles")


def test_air_coupled_batch_matches_alone():
    from worldinagrain_amd.gpu import GpuBatch
    worlds = spec() + [("still tank 48x40", 48, 40, scenes.still_tank(48, 40, depth=20))]
    batched, alone = [], []
    for i, (name, W, H, cells) in enumerate(worlds):
        wb, wa = make_gpu(W, H, cells), make_gpu(W, H, cells)
        if i < 2:
            batched.append((wb, coupled(wb, W, H, cells, 11 + i)))
            alone.append((wa, coupled(wa, W, H, cells, 11 + i)))
        else:                                   # the air not coupled: at rest, airSource untouched
            batched.append((wb, config_sim(wb, W, H, tol=1e-6)))
            alone.append((wa, config_sim(wa, W, H, tol=1e-6)))
    batch = GpuBatch([w for w, _ in batched])
    for s in range(5):
        say(f"[air coupling batch] step {s}")
        step_batch(batch, [sim for _, sim in batched])
        for _, sa in alone:
            sa.step()
        for (name, *_), (wb, sb), (wa, sa) in zip(worlds, batched, alone):
            assert np.array_equal(sb.last_stats, sa.last_stats), (name, s)
            d = compare_worlds(wb, wa)
            assert not d, f"{name} step {s}: " + "; ".join(d)
            same_source(wb, wa, f"{name} step {s}")
    idle = np.asarray(batched[2][0].download(abi.PART_AIR_SOURCE)).view(np.uint32)
    assert not idle.any()
