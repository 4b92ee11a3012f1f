"""The host side of the liquid step: what `Simulation` in js/sim.js keeps and
decides around `encodeFluid` (js/sim.js:1386-1397,1893-1901,1972-2001,
2466-2612), over any world that takes (params, ctl) steps: the HIP world of
`worldinagrain_amd.gpu` in the product.

Per step it writes the Params slot (size, frame, a mulberry32 seed, dx,
earthG, dt), flips the relaxation's starting tiling, and takes the iteration
budget and the active/idle decision from the last report, as the reference
host does from its readback (here the report is read right after each step:
the reference's lands one or two frames late)."""
from dataclasses import dataclass

import numpy as np

from . import abi
from .elements import GRAVITY

FLUID_CYCLES = 16          # js/sim.js:63
FLUID_CYCLES_SPARE = 2     # js/sim.js:70
LIQUID_TOLERANCE = 1e-2    # fluid.wgsl:2221


@dataclass
class SolverSettings:
    """The pressure solve's settings. `reference()` is the reference's own
    (js/sim.js:61-70; fluid.wgsl:2063,2221,2392)."""
    tol_mode: int = abi.WIG_TOL_ABS_MAX
    tol: float = LIQUID_TOLERANCE
    max_cycles: int = FLUID_CYCLES
    adaptive_cycles: bool = True
    pre_relax: int = 1
    post_relax: int = 1
    sweep_pairs: int = 4
    coarse_sweeps: int = 40
    # this build's: iterations encoded before the first convergence check (0:
    # all fluid_cycles, the reference's way; -1: as many as the last step
    # ran), and then how many more at a time
    first_chunk: int = 0
    chunk: int = 2
    level_pairs: int = 1     # red-black pairs a coarse level, down and up (the reference's 1)
    storage: int = abi.WIG_STORE_F32   # the V-cycle's storage (WIG_STORE_*)
    # this build's: every k-th iteration the residual recomputed as b - A x
    # and the stop judged only there (0: the recursive residual throughout)
    residual_every: int = 0

    @staticmethod
    def reference():
        return SolverSettings()

    @staticmethod
    def relative(tol, max_cycles=100, sweep_pairs=2, level_pairs=2, coarse_sweeps=10, storage=abi.WIG_STORE_F32,
                 residual_every=0):
        """A solve to ||r|| <= tol ||b|| (the benchmark configurations: 2
        red-black sweep pairs before and after the coarse levels, and on each
        of them; 10 SOR sweep pairs on the coarsest, which converge the
        iteration as fast as the reference's 40)."""
        return SolverSettings(tol_mode=abi.WIG_TOL_REL_L2, tol=tol, max_cycles=max_cycles, adaptive_cycles=False,
                              sweep_pairs=sweep_pairs, first_chunk=-1, level_pairs=level_pairs,
                              coarse_sweeps=coarse_sweeps, storage=storage, residual_every=residual_every)


KARMAN = np.float32(0.41)               # js/shared.js:289
SURFACE_ROUGHNESS = np.float32(1e-4)    # fluid.wgsl:1890


CAPILLARY_LENGTH = 0.0027        # js/shared.js:285-286, water's, metres
SURFACE_MIN_LENGTH = 0.3


def surface_tension(dx):
    """Whether a cell of dx metres is small enough for surface tension to
    show (Simulation.surfaceTension, js/sim.js:2449-2454)."""
    return CAPILLARY_LENGTH / dx >= SURFACE_MIN_LENGTH


def surface_drag_coefficient(dx):
    """The wind's drag coefficient along a liquid surface, (KARMAN /
    ln(max(dx / 2 / z0, 2)))^2 in f32 (fluid.wgsl:1884-1895), worked out on
    the host once a step so every backend takes the same float (the log in
    double, rounded to f32, as the JavaScript host does)."""
    x = max(np.float32(np.float32(0.5) * np.float32(dx)) / SURFACE_ROUGHNESS, np.float32(2.0))
    r = KARMAN / np.float32(np.log(np.float64(np.float32(x))))
    return float(np.float32(r * r))


def mulberry32_next(state):
    """(new state, u32): js/sim.js:1390-1397 (Tommy Ettinger's mulberry32)."""
    state = (state + 0x6D2B79F5) & 0xFFFFFFFF
    t = state
    t = ((t ^ (t >> 15)) * (t | 1)) & 0xFFFFFFFF
    t ^= (t + (((t ^ (t >> 7)) * (t | 61)) & 0xFFFFFFFF)) & 0xFFFFFFFF
    return state, (t ^ (t >> 14)) & 0xFFFFFFFF


class Simulation:
    """The liquid's part of js/sim.js's Simulation over a world backend."""

    def __init__(self, world, width, height, *, seed=0x5EED, dx=0.01, earth_g=None, dt=1.0 / 60.0,
                 solver=None, frame=1):
        self.world = world
        self.width, self.height = width, height
        self.rng_state = seed & 0xFFFFFFFF
        self.dx = dx
        self.earth_g = float(GRAVITY) if earth_g is None else earth_g
        self.dt = dt
        self.solver = solver or SolverSettings.reference()
        self.frame_no = frame
        self.shifted_first = False
        self.fluid_stats = {"active": 1, "viscous": 0}
        self.fluid_cycles = self.solver.max_cycles
        self.last_stats = None
        self.last_cycles = 0
        # the air coupled in (WigStepCtl.air): the world's air part read, its
        # airSource part added to; off, the air is at rest
        self.air = False
        # people's bodies (WigStepCtl.occupants): the world's occupants part
        # read as moving obstacles; off, nobody is there
        self.occupants = False
        # the gravity pass's field (WigStepCtl.gravity): the world's gravity
        # part read for the pull and down; off, Earth's pull alone
        self.gravity = False
        # free and held bodies (WigStepCtl.bodies): the world's bodyMotion and
        # bodySwept parts read, bodyPushes its scratch; off, no body anywhere.
        # bodies_run is the reference's bodiesRun() (something loose as of the
        # last report, or that report too old: js/sim.js:2375-2383), which the
        # caller's body pass knows: it makes the step full and hands what the
        # liquid did to the bodies (rigidify, bodyPush)
        self.bodies = False
        self.bodies_run = False

    # Simulation.seed / random (js/sim.js:1386-1397)
    def seed(self, n):
        self.rng_state = n & 0xFFFFFFFF

    def random(self):
        self.rng_state, r = mulberry32_next(self.rng_state)
        return r

    def params(self):
        """writeSlot (js/sim.js:1972-2001), the words the liquid reads."""
        p = abi.WigParams()
        p.size[0], p.size[1] = self.width, self.height
        p.frame = self.frame_no & 0xFFFFFFFF
        p.seed = self.random()
        p.dx = self.dx
        p.earthG = self.earth_g
        p.dt = self.dt
        return p

    def ctl(self):
        s = self.solver
        self.shifted_first = not self.shifted_first       # js/sim.js:2530
        c = abi.WigStepCtl()
        run = bool(self.bodies and self.bodies_run)
        c.full = 1 if self.fluid_stats["active"] or run else 0   # js/sim.js:2479
        c.bodies_run = 1 if run else 0
        c.bodies = 1 if self.bodies else 0
        # the surface pass, and viscosity's sweeps with it (js/sim.js:2496-2509)
        c.surface = 1 if surface_tension(self.dx) else 0
        c.viscous = 1 if self.fluid_stats["viscous"] or c.surface else 0
        c.fluid_cycles = self.fluid_cycles if s.adaptive_cycles else s.max_cycles
        c.shifted_first = 1 if self.shifted_first else 0
        c.tol_mode = s.tol_mode
        c.tol = s.tol
        c.pre_relax, c.post_relax = s.pre_relax, s.post_relax
        c.sweep_pairs, c.coarse_sweeps = s.sweep_pairs, s.coarse_sweeps
        c.level_pairs = s.level_pairs
        c.vcycle_storage = s.storage
        c.residual_every = s.residual_every
        c.air = 1 if self.air else 0
        c.surface_drag = surface_drag_coefficient(self.dx)
        c.occupants = 1 if self.occupants else 0
        c.gravity = 1 if self.gravity else 0
        if s.first_chunk < 0:
            c.first_chunk = 0xFFFFFFFF          # WIG_CHUNK_AUTO: the world's last solve's count
            c.chunk = s.chunk
        else:
            c.first_chunk, c.chunk = s.first_chunk, s.chunk
        return c

    def step(self, read_report=True):
        """One step's liquid (encodeFluid), then its report taken in.
        read_report "late": the report asked for without waiting and the
        last step's taken in instead, once landed, as the reference takes a
        report in when its mapAsync resolves (js/sim.js:1882-1901): the
        host never waits for the step it has just enqueued."""
        status = self.world.step(self.params(), self.ctl())
        if status < 0:
            raise RuntimeError(f"fluid step failed ({status})")
        self.frame_no += 1
        if read_report == "late":
            last, self.pending_report = getattr(self, "pending_report", None), self.world.report_async()
            if last is not None:
                self.take_report(self.world.report_take(last, wait=True))
        elif read_report:
            self.take_report(self.world.report())
        return status

    def take_report(self, stats):
        """readReport's liquid part (js/sim.js:1893-1901)."""
        stats = np.asarray(stats, dtype=np.uint32)
        self.last_stats = stats
        self.fluid_stats = {"active": int(stats[abi.STAT_ACTIVE]) + int(stats[abi.STAT_PARTICLES]),
                            "viscous": int(stats[abi.STAT_VISCOUS])}
        self.last_cycles = int(stats[abi.STAT_CYCLES])
        most = int(stats[abi.STAT_MOST_CYCLES])
        cap = self.solver.max_cycles
        self.fluid_cycles = cap if most >= self.fluid_cycles else min(most + FLUID_CYCLES_SPARE, cap)


def step_batch(batch, sims, read_report=True):
    """One step of each simulation, all in one batch step (wig_batch_step):
    each takes its own Params and controls as its step would, and its own
    report after."""
    params = [s.params() for s in sims]
    ctls = [s.ctl() for s in sims]
    status = batch.step(params, ctls)
    for s in sims:
        s.frame_no += 1
    if read_report:
        for s, row in zip(sims, batch.report()):
            s.take_report(row)
    return status


def mulberry32_many(state):
    """mulberry32_next over an array of uint32 states: (new states, u32s)."""
    st = (state.astype(np.uint64) + 0x6D2B79F5) & 0xFFFFFFFF
    t = st.copy()
    t = ((t ^ (t >> 15)) * (t | 1)) & 0xFFFFFFFF
    t ^= (t + (((t ^ (t >> 7)) * (t | 61)) & 0xFFFFFFFF)) & 0xFFFFFFFF
    return st.astype(np.uint32), ((t ^ (t >> 14)) & 0xFFFFFFFF).astype(np.uint32)


class BatchHost:
    """The host side of many Simulations stepped as one batch: each step's
    Params and controls for all of them at once (numpy), the same words each
    Simulation.step would write, and their reports taken in after.  The
    simulations' own state (seed, frame, tiling flip, report) is kept up to
    date, so they can be stepped alone again at any time."""

    def __init__(self, sims):
        self.sims = list(sims)
        n = len(self.sims)
        self.P = (abi.WigParams * n)()
        self.C = (abi.WigStepCtl * n)()
        for i, s in enumerate(self.sims):
            p = self.P[i]
            p.size[0], p.size[1] = s.width, s.height
            p.dx, p.earthG, p.dt = s.dx, s.earth_g, s.dt
            c = self.C[i]
            sv = s.solver
            c.tol_mode, c.tol = sv.tol_mode, sv.tol
            c.pre_relax, c.post_relax = sv.pre_relax, sv.post_relax
            c.sweep_pairs, c.coarse_sweeps = sv.sweep_pairs, sv.coarse_sweeps
            c.level_pairs = sv.level_pairs
            c.vcycle_storage = sv.storage
            c.residual_every = sv.residual_every
            c.air = 1 if s.air else 0
            c.surface_drag = surface_drag_coefficient(s.dx)
            c.occupants = 1 if s.occupants else 0
            c.gravity = 1 if s.gravity else 0
            c.bodies = 1 if s.bodies else 0
            c.bodies_run = 1 if s.bodies and s.bodies_run else 0
            if sv.first_chunk < 0:
                c.first_chunk, c.chunk = 0xFFFFFFFF, sv.chunk
            else:
                c.first_chunk, c.chunk = sv.first_chunk, sv.chunk
        self.pw = np.frombuffer(self.P, dtype=np.uint32).reshape(n, -1)
        self.cw = np.frombuffer(self.C, dtype=np.uint32).reshape(n, -1)
        self._frame = abi.WigParams.frame.offset // 4
        self._seed = abi.WigParams.seed.offset // 4
        self._full = abi.WigStepCtl.full.offset // 4
        self._visc = abi.WigStepCtl.viscous.offset // 4
        self._surf = np.array([1 if surface_tension(s.dx) else 0 for s in self.sims], dtype=np.uint32)
        self.cw[:, abi.WigStepCtl.surface.offset // 4] = self._surf
        self._cycles = abi.WigStepCtl.fluid_cycles.offset // 4
        self._shift = abi.WigStepCtl.shifted_first.offset // 4
        self.adaptive = np.array([s.solver.adaptive_cycles for s in self.sims])
        self.cap = np.array([s.solver.max_cycles for s in self.sims], dtype=np.uint32)

    def step(self, batch, read_report=True):
        sims = self.sims
        rng = np.array([s.rng_state for s in sims], dtype=np.uint32)
        rng, seeds = mulberry32_many(rng)
        shifted = np.array([not s.shifted_first for s in sims])
        pw, cw = self.pw, self.cw
        pw[:, self._frame] = np.array([s.frame_no & 0xFFFFFFFF for s in sims], dtype=np.uint32)
        pw[:, self._seed] = seeds
        cw[:, self._full] = np.array([1 if s.fluid_stats["active"] or (s.bodies and s.bodies_run) else 0 for s in sims],
                                     dtype=np.uint32)
        cw[:, self._visc] = np.array([1 if s.fluid_stats["viscous"] else 0 for s in sims], dtype=np.uint32) | self._surf
        cyc = np.array([s.fluid_cycles for s in sims], dtype=np.uint32)
        cw[:, self._cycles] = np.where(self.adaptive, cyc, self.cap)
        cw[:, self._shift] = shifted.astype(np.uint32)
        status = batch.step_arrays(self.P, self.C)
        for s, r, sh in zip(sims, rng.tolist(), shifted.tolist()):
            s.rng_state = r
            s.shifted_first = sh
            s.frame_no += 1
        if read_report:
            rows = batch.report()
            active = rows[:, abi.STAT_ACTIVE].astype(np.int64) + rows[:, abi.STAT_PARTICLES]
            most = rows[:, abi.STAT_MOST_CYCLES]
            cap = self.cap
            new = np.where(most >= cyc, cap, np.minimum(most + FLUID_CYCLES_SPARE, cap))
            for i, s in enumerate(sims):
                s.last_stats = rows[i]
                s.fluid_stats = {"active": int(active[i]), "viscous": int(rows[i, abi.STAT_VISCOUS])}
                s.last_cycles = int(rows[i, abi.STAT_CYCLES])
                s.fluid_cycles = int(new[i])
        return status
