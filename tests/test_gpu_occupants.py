"""People's bodies in the liquid step (WigStepCtl.occupants) on the GPU
against the oracle: a person, a rectangle of occupied cells moving through
a dam break, stamped anew in the occupants part before each step as the
reference's people pass does; every buffer bit for bit.  Then the same world
in a batch beside one with nobody in it, each as it steps alone."""
import sys

import numpy as np
import pytest

from worldinagrain_amd import abi, scenes
from worldinagrain_amd.host import step_batch
from tests.helpers import make_oracle, make_gpu, compare_worlds, config_sim

pytestmark = pytest.mark.gpu

W, H = 96, 80


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def person(step, vx=0.75, vy=-0.25):
    """The occupants part at `step`: a 5 x 12 body walking right through the
    water at about vx cells a step, its velocity in every word."""
    occ = np.zeros((H, W), dtype=np.uint32)
    x0 = 20 + int(round(vx * step))
    occ[H - 14:H - 2, x0:x0 + 5] = abi.occupant_word(4, vx, vy)
    return occ.reshape(-1)


def stepped(world):
    s = config_sim(world, W, H, tol=1e-6)
    s.occupants = True
    return s


def test_occupants_match_oracle():
    cells = scenes.dam_break(W, H, block=(60, 50))
    wo, wg = make_oracle(W, H, cells), make_gpu(W, H, cells)
    so, sg = stepped(wo), stepped(wg)
    for s in range(10):
        occ = person(s)
        wo.upload(abi.PART_OCCUPANTS, occ)
        wg.upload(abi.PART_OCCUPANTS, occ)
        so.step()
        sg.step()
        assert np.array_equal(so.last_stats, sg.last_stats), (s, so.last_stats, sg.last_stats)
        d = compare_worlds(wo, wg)
        assert not d, f"step {s}: " + "; ".join(d)
        # the person's cells hold no drawn liquid after raster
        words = np.asarray(wg.download(abi.PART_CELLS))[:4 * W * H].reshape(H, W, 4)
        held = occ.reshape(H, W) != 0
        assert not np.any((words[..., 0][held] & 0xFF) == 3), f"step {s}: liquid drawn under the person"
    say(f"[occupants] 10 steps, GPU == oracle; {int(sg.last_stats[abi.STAT_PARTICLES])} particles")


def test_occupants_batch_matches_alone():
    from worldinagrain_amd.gpu import GpuBatch
    cells = scenes.dam_break(W, H, block=(60, 50))
    pairs = []
    for with_person in (True, False):
        wb, wa = make_gpu(W, H, cells), make_gpu(W, H, cells)
        sb, sa = (stepped(wb), stepped(wa)) if with_person else (config_sim(wb, W, H, tol=1e-6), config_sim(wa, W, H, tol=1e-6))
        pairs.append((with_person, wb, sb, wa, sa))
    batch = GpuBatch([p[1] for p in pairs])
    for s in range(5):
        for with_person, wb, _, wa, _ in pairs:
            if with_person:
                wb.upload(abi.PART_OCCUPANTS, person(s))
                wa.upload(abi.PART_OCCUPANTS, person(s))
        step_batch(batch, [p[2] for p in pairs])
        for with_person, wb, sb, wa, sa in pairs:
            sa.step()
            assert np.array_equal(sb.last_stats, sa.last_stats), (with_person, s)
            d = compare_worlds(wb, wa)
            assert not d, f"person {with_person} step {s}: " + "; ".join(d)
