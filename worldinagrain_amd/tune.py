"""The autotuner: the coarse-level sweep kernels' tile sides and the
smoother's sweep counts, chosen by timing them on the world they are for.

Tile sides do not change results (each tile's halo makes its sweeps equal
the whole-level passes), so they are tuned on the world's own steps.  Sweep
counts change the iteration and its results, so each candidate runs on a
copy of the world (its parts copied into a scratch world) from the same
state, and is scored by the solve's time a step at the solver's own
tolerance; the world itself is not stepped."""
import copy
import dataclasses

from .gpu import WigError

TILE_SIDES = (8, 16, 32)
SWEEP_CANDIDATES = [(sp, lp, cs) for sp in (1, 2, 3) for lp in (1, 2, 3) for cs in (1, 2, 5)]


def _kernel_ms_per_iteration(world, sim, name, steps):
    t0 = world.run_totals()
    world.time_kernel(name)
    for _ in range(steps):
        sim.step()
    ms, _ = world.kernel_time()
    world.time_kernel(None)
    cycles = world.run_totals().cycles - t0.cycles
    return ms / max(cycles, 1)


def tune_tiles(world, sim, sides=TILE_SIDES, steps=2):
    """Each tiled coarse level's side in turn, by the tile kernels' time an
    iteration over `steps` of the world's steps.  Returns {level: side}."""
    chosen = {}
    level = 1
    while True:
        try:
            world.set_tile(level, sides[0])
        except WigError:
            break       # past the tiled levels
        best = None
        for side in sides:
            world.set_tile(level, side)
            ms = sum(_kernel_ms_per_iteration(world, sim, k, steps) for k in ("levelDown", "levelUp"))
            if best is None or ms < best[0]:
                best = (ms, side)
        world.set_tile(level, best[1])
        chosen[level] = best[1]
        level += 1
    return chosen


WAVE_CELLS = (0, 256, 1024)


def tune_lds_level(world, sim, steps=3, waves=WAVE_CELLS):
    """Where the one-workgroup coarse-level kernel starts, and from which
    level down its first wavefront works alone: each level from the first
    LDS holds down to the coarsest with each of `waves` (wig_set_coarse_wave),
    by the solve's time an iteration over `steps` of the world's steps
    (results depend on neither).  Returns (level, wave cells)."""
    best = None
    for cand in range(1, world.layout.levels + 1):
        try:
            world.set_lds_level(cand)
        except WigError:
            continue
        for cells in waves:
            world.set_coarse_wave(cells)
            ms = _kernel_ms_per_iteration(world, sim, "solve", steps)
            if best is None or ms < best[0]:
                best = (ms, cand, cells)
    if best is None:
        return None, None
    world.set_lds_level(best[1])
    world.set_coarse_wave(best[2])
    return best[1], best[2]


def tune_coarse_wave_batch(batch, sims, groups, waves=WAVE_CELLS, steps=1):
    """The one-wavefront coarse levels for worlds stepped in a batch: for
    each group of worlds (indices into the batch), each of `waves` in turn
    for the whole group, scored by the batch's coarse-level kernel's mean
    launch time over `steps` batch steps.  Returns {group key: cells}."""
    from .host import step_batch

    def score():
        batch.time_kernel("coarseLDS")
        for _ in range(steps):
            step_batch(batch, sims, read_report=False)
        t, n = batch.kernel_time()
        batch.time_kernel(None)
        return t / max(n, 1)

    chosen = {}
    for key, idx in groups.items():
        best = None
        for cells in waves:
            for i in idx:
                batch.worlds[i].set_coarse_wave(cells)
            ms = score()
            if best is None or ms < best[0]:
                best = (ms, cells)
        for i in idx:
            batch.worlds[i].set_coarse_wave(best[1])
        chosen[key] = best[1]
    return chosen


def tune_tiles_batch(batch, sims, groups, sides=TILE_SIDES, steps=1):
    """Tile sides for worlds stepped in a batch, where the tile kernels
    share the device with every other world's: for each group of worlds
    (a list of indices into the batch, e.g. one side class) and each tiled
    level, each side in turn for the whole group, scored by the batch's
    tile kernels' mean launch time over `steps` batch steps.  Returns
    {group key: {level: side}}."""
    from .host import step_batch

    def score():
        ms = 0.0
        for k in ("levelDown", "levelUp"):
            batch.time_kernel(k)
            for _ in range(steps):
                step_batch(batch, sims, read_report=False)
            t, n = batch.kernel_time()
            ms += t / max(n, 1)
        batch.time_kernel(None)
        return ms

    chosen = {}
    for key, idx in groups.items():
        chosen[key] = {}
        worlds = [batch.worlds[i] for i in idx]
        level = 1
        while True:
            try:
                for w in worlds:
                    w.set_tile(level, sides[0])
            except WigError:
                break
            best = None
            for side in sides:
                for w in worlds:
                    w.set_tile(level, side)
                ms = score()
                if best is None or ms < best[0]:
                    best = (ms, side)
            for w in worlds:
                w.set_tile(level, best[1])
            chosen[key][level] = best[1]
            level += 1
    return chosen


def tune_sweeps(world, sim, candidates=SWEEP_CANDIDATES, steps=3):
    """(sweep_pairs, level_pairs, coarse_sweeps) by the solve's time a step
    on copies of the world (ties within 3 % to fewer iterations).  Returns (best settings, [(ms a step, iterations
    a step, candidate), ...]); `sim.solver` is left as it was."""
    scores = []
    for sp, lp, cs in candidates:
        w = world.clone()
        try:
            s = copy.copy(sim)
            s.world = w
            s.fluid_stats = dict(sim.fluid_stats)
            s.solver = dataclasses.replace(sim.solver, sweep_pairs=sp, level_pairs=lp, coarse_sweeps=cs)
            s.step()            # its first step finds the iteration count the chunks start from
            t0 = w.run_totals()
            w.time_kernel("solve")
            for _ in range(steps):
                s.step()
            ms, _ = w.kernel_time()
            w.time_kernel(None)
            its = (w.run_totals().cycles - t0.cycles) / steps
            scores.append((ms / steps, its, (sp, lp, cs)))
        finally:
            w.close()
    # the fewest iterations among those within 3 % of the fastest: a
    # world alone is bound by its passes' latency, and the iteration count
    # is what counts when it steps beside others
    fastest = min(scores)[0]
    best = min((its, ms, cand) for ms, its, cand in scores if ms <= 1.03 * fastest)[2]
    return dataclasses.replace(sim.solver, sweep_pairs=best[0], level_pairs=best[1], coarse_sweeps=best[2]), scores
