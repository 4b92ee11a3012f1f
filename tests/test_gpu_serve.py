"""The serving loop (worldinagrain_amd.serve): requests of many sessions,
continuous batching under an HBM budget small enough to force evictions.
Each session's world must end bit-identical to the same world stepped alone
as many times, whatever it was batched with and however often it was
evicted to host memory and admitted again."""
import sys

import numpy as np
import pytest

from worldinagrain_amd import abi, scenes, elements as E
from worldinagrain_amd.host import SolverSettings
from worldinagrain_amd.gpu import HipStream
from worldinagrain_amd.serve import Server, Session, world_bytes
from tests.helpers import make_gpu, compare_worlds, config_sim

pytestmark = pytest.mark.gpu

SIM_KW = dict(dx=1.0, earth_g=float(np.float32(9.81 * (1 / 120) ** 2)), dt=1 / 120)


@pytest.mark.parametrize("evictor", [False, True])
def test_served_results_equal_alone(evictor):
    sides = [64, 96, 128, 64, 160, 96, 128, 64]
    sessions = []
    for i, side in enumerate(sides):
        cells, _ = scenes.batch_world(side, seed=300 + i)
        sessions.append(Session(i, side, side, cells, lambda: SolverSettings.relative(1e-6), E.kinds(), sim_kw=SIM_KW))
    budget = world_bytes(160, 160) + 2 * world_bytes(128, 128)     # about three worlds at a time
    rng = np.random.default_rng(5)
    arrivals = sorted((float(rng.uniform(0, 0.05)), int(rng.integers(len(sides))), int(rng.integers(1, 4))) for _ in range(24))
    # evictions in their thread, on a copy stream of their own, with at most
    # about a world in flight (the loop waits for the rest)
    copy = HipStream() if evictor else None
    print(f"[serve] evictor {evictor}", file=sys.stderr, flush=True)
    server = Server(sessions, budget, max_round_cells=160 * 160 * 3,
                    copy_stream=copy.handle if copy is not None else None, evicting_limit=world_bytes(160, 160))
    reqs = server.run(arrivals)
    print(f"[serve] rounds {server.round_no} evictions {server.evictions} admissions {server.admissions}", file=sys.stderr, flush=True)
    if server.evictor is not None:
        server.evictor.close()
        server.evictor = None
    assert all(r.done_at is not None for r in reqs)
    assert server.evictions > 0
    for s in sessions:
        if s.steps_run == 0:
            continue
        if s.world is None:
            s.make_resident()
        alone = make_gpu(s.width, s.height, s.cells)
        sim = config_sim(alone, s.width, s.height)
        for _ in range(s.steps_run):
            sim.step()
        d = compare_worlds(s.world, alone)
        assert not d, f"session {s.sid} ({s.steps_run} steps): " + "; ".join(d)
    server.close()
    if copy is not None:
        copy.close()
