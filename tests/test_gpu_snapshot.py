"""Snapshots of HIP worlds: a world saved mid-run and loaded into a fresh
one steps on exactly as the world never saved; the same snapshot loaded
into the oracle's world steps on to the same buffers (a snapshot does not
depend on the backend that wrote it)."""
import sys

import numpy as np
import pytest

from worldinagrain_amd import abi, scenes, snapshot
from worldinagrain_amd.host import Simulation
from tests.helpers import make_gpu, make_oracle, compare_worlds, config_sim

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("mode", ["reference", "config"])
def test_snapshot_continues_exactly(mode):
    W, H = 96, 80
    cells, _ = scenes.discs_dam(W, H, discs=8, rmin=3, rmax=7)
    say(f"[snapshot] {W}x{H} {mode}")
    make = Simulation if mode == "reference" else config_sim
    a = make(make_gpu(W, H, cells), W, H)
    for _ in range(4):
        a.step()
    blob = snapshot.save(a, level=1)
    b = make(make_gpu(W, H, cells), W, H)
    snapshot.load(b, blob)
    o = make(make_oracle(W, H, cells), W, H)
    snapshot.load(o, blob)
    for s in range(4):
        a.step()
        b.step()
        o.step()
        assert np.array_equal(a.last_stats, b.last_stats), s
        assert np.array_equal(a.last_stats, o.last_stats), s
    diffs = compare_worlds(a.world, b.world)
    assert not diffs, "; ".join(diffs)
    diffs = compare_worlds(a.world, o.world)
    assert not diffs, "; ".join(diffs)


@pytest.mark.parametrize("storage", [abi.WIG_STORE_F16, abi.WIG_STORE_BF16])
def test_snapshot_continues_exactly_16bit(storage):
    # the 16-bit V-cycle's state (its scale, the solve part's words) travels
    # in the snapshot: a world loaded from it steps on as the world saved
    from worldinagrain_amd.host import SolverSettings
    W, H = 96, 80
    cells, _ = scenes.discs_dam(W, H, discs=8, rmin=3, rmax=7)
    say(f"[snapshot] {W}x{H} storage {storage}")
    sims = []
    for _ in range(2):
        s = config_sim(make_gpu(W, H, cells), W, H)
        s.solver = SolverSettings.relative(1e-6, storage=storage)
        sims.append(s)
    a, b = sims
    for _ in range(4):
        a.step()
    snapshot.load(b, snapshot.save(a, level=1))
    for s in range(4):
        a.step()
        b.step()
        assert np.array_equal(a.last_stats, b.last_stats), s
    diffs = compare_worlds(a.world, b.world)
    assert not diffs, "; ".join(diffs)
