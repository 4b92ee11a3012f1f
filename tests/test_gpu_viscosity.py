"""Viscosity in the liquid step on the GPU against the oracle: a dam break of
a thick liquid beside water and a liquid whose viscosity falls with its
temperature (heated from one side), the sweeps run from the first report
that finds viscous liquid (js/sim.js:2505-2514).  Every buffer bit for bit,
viscosity's face words (the sweeps' right-hand sides, what they did last
step) included; then the same world in a batch beside plain water, each as
it steps alone."""
import sys

import numpy as np
import pytest

from worldinagrain_amd import abi, scenes, elements as E
from worldinagrain_amd.host import step_batch
from tests.helpers import compare_worlds, config_sim
from tests.oracle_pins import viscous_elements, VISCOUS, HEATED

pytestmark = pytest.mark.gpu

W, H = 96, 80


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def scene():
    cells = scenes.dam_break(W, H, block=(70, 40))
    words = cells[:4 * W * H].reshape(H, W, 4)
    liquid = (words[..., 0] & 0xFF) == E.WATER
    x = np.arange(W)[None, :].repeat(H, 0)
    thick, heated = liquid & (x < 24), liquid & (x >= 24) & (x < 48)
    words[thick, 0] = (words[thick, 0] & np.uint32(0xFFFFFF00)) | np.uint32(VISCOUS)
    words[heated, 0] = (words[heated, 0] & np.uint32(0xFFFFFF00)) | np.uint32(HEATED)
    temps = (20.0 + 2.5 * (x - 24)).astype(np.float32)
    words[heated, 1] = temps[heated].view(np.uint32)
    return cells


def world(kind, cells):
    kinds, props = viscous_elements()
    if kind == "oracle":
        from oracle.oracle import OracleWorld
        w = OracleWorld(W, H, kinds=kinds, dot_mode=1)
    else:
        from worldinagrain_amd.gpu import GpuWorld
        w = GpuWorld(W, H, kinds=kinds)
    w.set_element_props(props)
    w.upload(abi.PART_CELLS, cells)
    return w


def test_viscosity_matches_oracle():
    cells = scene()
    wo, wg = world("oracle", cells), world("gpu", cells)
    so, sg = config_sim(wo, W, H, tol=1e-6), config_sim(wg, W, H, tol=1e-6)
    for s in range(12):
        so.step()
        sg.step()
        assert np.array_equal(so.last_stats, sg.last_stats), (s, so.last_stats, sg.last_stats)
        d = compare_worlds(wo, wg)
        assert not d, f"step {s}: " + "; ".join(d)
    assert int(sg.last_stats[abi.STAT_VISCOUS]) == 1
    # the sweeps did something: viscosity's last-change words are not all zero
    f = wg.download(abi.PART_FIELD)
    C, F = W * H, (W + 1) * H + W * (H + 1)
    last = f[4 * (2 * C + F + (F + 3) // 4):][:F]
    assert np.any(last != 0.0)
    say(f"[viscosity] 12 steps, GPU == oracle; {int(sg.last_stats[abi.STAT_PARTICLES])} particles")


def test_viscosity_batch_matches_alone():
    from worldinagrain_amd.gpu import GpuBatch
    from tests.helpers import make_gpu
    cells = scene()
    plain = scenes.dam_break(W, H, block=(70, 40))
    pairs = []
    for c, mk in ((cells, lambda c: world("gpu", c)), (plain, lambda c: make_gpu(W, H, c))):
        wb, wa = mk(c), mk(c)
        pairs.append((wb, config_sim(wb, W, H, tol=1e-6), wa, config_sim(wa, W, H, tol=1e-6)))
    batch = GpuBatch([p[0] for p in pairs])
    for s in range(6):
        step_batch(batch, [p[1] for p in pairs])
        for wb, sb, wa, sa in pairs:
            sa.step()
            assert np.array_equal(sb.last_stats, sa.last_stats), s
            d = compare_worlds(wb, wa)
            assert not d, f"step {s}: " + "; ".join(d)
