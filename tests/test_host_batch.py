"""The batch host (host.BatchHost) writes, for each simulation of a batch,
the same Params and controls its own Simulation.step would, and takes in
its report the same way: checked on the CPU against simulations stepped one
by one, over steps with varied reports (no GPU: the batch is a recorder)."""
import copy

import numpy as np

from worldinagrain_amd import abi
from worldinagrain_amd.host import BatchHost, Simulation, SolverSettings


class RecordingWorld:
    def __init__(self, rows):
        self.rows, self.k, self.seen = rows, 0, []

    def step(self, params, ctl):
        self.seen.append((bytes(params), bytes(ctl)))
        return 0

    def report(self):
        r = self.rows[self.k]
        self.k += 1
        return r


class RecordingBatch:
    def __init__(self, rows):
        self.rows, self.k, self.seen = rows, 0, []

    def step_arrays(self, P, C):
        self.seen.append([(bytes(P[i]), bytes(C[i])) for i in range(len(P))])
        return 0

    def report(self):
        r = np.stack([rw[self.k] for rw in self.rows])
        self.k += 1
        return r


def test_batch_host_matches_simulations():
    rng = np.random.default_rng(3)
    n, steps = 7, 6
    rows = [rng.integers(0, 40, size=(steps, abi.STATS)).astype(np.uint32) for _ in range(n)]
    for r in rows:
        r[2, abi.STAT_ACTIVE] = 0
        r[2, abi.STAT_PARTICLES] = 0          # an idle report: the next step takes the idle path
    solvers = [SolverSettings.reference(), SolverSettings.relative(1e-6), SolverSettings.relative(1e-5, storage=abi.WIG_STORE_F16)]
    alone, batched = [], []
    for i in range(n):
        w = RecordingWorld(rows[i])
        s = Simulation(w, 32 + 8 * i, 24 + 8 * i, seed=1000 + i, dx=1.0, earth_g=1e-3 * i, dt=1 / 120,
                       solver=copy.deepcopy(solvers[i % 3]), frame=5 + i)
        alone.append(s)
        batched.append(Simulation(None, s.width, s.height, seed=1000 + i, dx=1.0, earth_g=1e-3 * i, dt=1 / 120,
                                  solver=copy.deepcopy(solvers[i % 3]), frame=5 + i))
    batch = RecordingBatch(rows)
    host = BatchHost(batched)
    for k in range(steps):
        for s in alone:
            s.step()
        host.step(batch)
        for i in range(n):
            assert batch.seen[k][i] == alone[i].world.seen[k], (k, i)
    for a, b in zip(alone, batched):
        assert (a.rng_state, a.frame_no, a.shifted_first, a.fluid_cycles, a.last_cycles) == \
               (b.rng_state, b.frame_no, b.shifted_first, b.fluid_cycles, b.last_cycles)
        assert a.fluid_stats == b.fluid_stats
        assert np.array_equal(a.last_stats, b.last_stats)
