"""Worlds stepped together in one batch (wig_batch_step) against the same
worlds stepped alone and against the oracle: each world's every buffer must
be bit-identical whatever batch it runs in.  The batch mixes sizes (some
with coarse levels swept in tiles), the reference's solver settings and the
relative-residual ones, and an idle world."""
import copy
import sys

import numpy as np
import pytest

from worldinagrain_amd import abi, scenes
from worldinagrain_amd.host import Simulation, step_batch
from tests.helpers import make_oracle, make_gpu, compare_worlds, config_sim

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def worlds_spec():
    discs, _ = scenes.discs_dam(330, 250, discs=40, rmin=3, rmax=12)
    small, _ = scenes.discs_dam(128, 96, discs=10, rmin=3, rmax=8)
    return [
        ("dam 64x48 reference", 64, 48, scenes.dam_break(64, 48, block=(30, 36)), "reference"),
        ("odd 61x37 reference", 61, 37, scenes.dam_break(61, 37, block=(29, 30)), "reference"),
        ("discs 330x250 relative", 330, 250, discs, "config"),
        ("idle 32x32", 32, 32, scenes.empty_cells(32, 32), "reference"),
        ("discs 128x96 relative", 128, 96, small, "config"),
    ]


def make_sim(world, W, H, mode):
    return Simulation(world, W, H) if mode == "reference" else config_sim(world, W, H)


def test_batch_matches_alone_and_oracle():
    from worldinagrain_amd.gpu import GpuBatch
    spec = worlds_spec()
    batched, alone, oracles = [], [], []
    for name, W, H, cells, mode in spec:
        wb, wa, wo = make_gpu(W, H, cells), make_gpu(W, H, cells), make_oracle(W, H, cells)
        batched.append((wb, make_sim(wb, W, H, mode)))
        alone.append((wa, make_sim(wa, W, H, mode)))
        oracles.append((wo, make_sim(wo, W, H, mode)))
    batch = GpuBatch([w for w, _ in batched])
    steps = 6
    for s in range(steps):
        say(f"[batch] step {s}")
        step_batch(batch, [sim for _, sim in batched])
        for (wa, sa), (wo, so) in zip(alone, oracles):
            sa.step()
            so.step()
        for i, (name, *_rest) in enumerate(spec):
            sb, sa, so = batched[i][1], alone[i][1], oracles[i][1]
            assert np.array_equal(sb.last_stats, sa.last_stats), (name, s, sb.last_stats, sa.last_stats)
            assert np.array_equal(sb.last_stats, so.last_stats), (name, s, sb.last_stats, so.last_stats)
    for i, (name, *_rest) in enumerate(spec):
        d = compare_worlds(batched[i][0], alone[i][0])
        assert not d, f"{name} batch vs alone: " + "; ".join(d)
        d = compare_worlds(batched[i][0], oracles[i][0])
        assert not d, f"{name} batch vs oracle: " + "; ".join(d)
    batch.close()


def test_batch_of_one_and_regrouping():
    """A world moves between batches of different company and still steps
    as it would alone."""
    from worldinagrain_amd.gpu import GpuBatch
    W, H = 96, 80
    cells, _ = scenes.discs_dam(W, H, discs=8, rmin=3, rmax=8)
    wa, wb = make_gpu(W, H, cells), make_gpu(W, H, cells)
    other = make_gpu(64, 48, scenes.dam_break(64, 48, block=(30, 36)))
    sa, sb, so = config_sim(wa, W, H), config_sim(wb, W, H), Simulation(other, 64, 48)
    b1 = GpuBatch([wb])
    b2 = GpuBatch([other, wb])
    for s in range(6):
        sa.step()
        if s % 2 == 0:
            step_batch(b1, [sb])
        else:
            step_batch(b2, [so, sb])
    d = compare_worlds(wa, wb)
    assert not d, "; ".join(d)


def test_graph_replays_match_eager():
    """Steps replayed as hipGraphs (a batch's and a lone world's) against the
    same worlds stepped with plain launches, every step: each replay runs
    after the host's own reads of the last step's reports."""
    from worldinagrain_amd.gpu import GpuBatch
    spec = worlds_spec()[1:4]
    eager, graphed = [], []
    for name, W, H, cells, mode in spec:
        we, wg = make_gpu(W, H, cells), make_gpu(W, H, cells)
        eager.append((name, we, make_sim(we, W, H, mode)))
        graphed.append((name, wg, make_sim(wg, W, H, mode)))
    batch = GpuBatch([w for _, w, _ in graphed])
    batch.set_graphs(True)
    W, H = 96, 80
    cells, _ = scenes.discs_dam(W, H, discs=8, rmin=3, rmax=8)
    lone_e, lone_g = make_gpu(W, H, cells), make_gpu(W, H, cells)
    lone_g.set_graphs(True)
    sle, slg = config_sim(lone_e, W, H), config_sim(lone_g, W, H)
    for s in range(5):
        say(f"[graphs] step {s}")
        step_batch(batch, [sim for _, _, sim in graphed])
        for _, _, sim in eager:
            sim.step()
        sle.step()
        slg.step()
        for (name, we, se), (_, wg, sg) in zip(eager, graphed):
            assert np.array_equal(se.last_stats, sg.last_stats), (name, s)
            d = compare_worlds(we, wg)
            assert not d, f"{name} step {s}: " + "; ".join(d)
        d = compare_worlds(lone_e, lone_g)
        assert not d, f"lone world step {s}: " + "; ".join(d)
    batch.close()


def test_batch_outlives_its_first_world():
    # a batch made without a stream runs on one of its own: its first world
    # assigned away and destroyed, it still steps the others exactly
    from worldinagrain_amd.gpu import GpuBatch
    spec = worlds_spec()
    say("[batch] first world destroyed")
    a = make_gpu(spec[0][1], spec[0][2], spec[0][3])
    b = make_gpu(spec[4][1], spec[4][2], spec[4][3])
    sa, sb = make_sim(a, spec[0][1], spec[0][2], "reference"), make_sim(b, spec[4][1], spec[4][2], "config")
    batch = GpuBatch([a, b])
    step_batch(batch, [sa, sb])
    batch.assign([b])
    a.close()
    for _ in range(2):
        step_batch(batch, [sb])
    alone = make_gpu(spec[4][1], spec[4][2], spec[4][3])
    s2 = make_sim(alone, spec[4][1], spec[4][2], "config")
    for _ in range(3):
        s2.step()
    d = compare_worlds(b, alone)
    assert not d, "; ".join(d)
    batch.close()
