"""Free and held bodies in the liquid step on the GPU against the oracle
(fluid.wgsl:1181-1475,2816-2845,2944-3218): a tank with a light plank
afloat, a heavy box sinking and turning under its contacts' mobility, a
held block on the floor and a cup with powder in it and a seam in its wall
(scenes.bodies_tank).  Every buffer bit for bit after every step, the
bodies' liquid slots and the step's push sums included; particles pushed
aside by a body (DEST_ASIDE), some toward a cup; with viscosity (the drag
on the bodies and their faces after the sweeps); and in a batch with
graphs, each world as it steps alone."""
import sys

import numpy as np
import pytest

from worldinagrain_amd import abi, scenes, elements as E
from worldinagrain_amd.host import Simulation, SolverSettings, step_batch
from tests.helpers import compare_worlds, first_difference

pytestmark = pytest.mark.gpu

W, H = 64, 48


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def world(kind, cells, motion, swept, props, kinds):
    if kind == "oracle":
        from oracle.oracle import OracleWorld
        w = OracleWorld(W, H, kinds=kinds, dot_mode=1)
    else:
        from worldinagrain_amd.gpu import GpuWorld
        w = GpuWorld(W, H, kinds=kinds)
    w.set_element_props(props)
    w.upload(abi.PART_CELLS, cells)
    w.upload(abi.PART_MOTION, motion)
    w.upload(abi.PART_SWEPT, swept)
    return w


def body_diffs(wo, wg):
    diffs = compare_worlds(wo, wg)
    for p in (abi.PART_MOTION, abi.PART_PUSHES):
        d = first_difference(np.asarray(wo.download(p)).view(np.uint32), np.asarray(wg.download(p)).view(np.uint32),
                             abi.PART_NAMES[p])
        if d:
            diffs.append(d)
    return diffs


def make_sim(w, solver):
    s = Simulation(w, W, H, solver=solver)
    s.bodies = s.bodies_run = True
    return s


def push_aside(worlds, frame):
    """Particles of a few cells pushed aside by a body between steps, as the
    bodies' pass marks them (G_DEST's DEST_ASIDE | the cell they go to): two
    toward cells in the cup's hold (kept where they are: cupOf), two toward
    open water."""
    for w in worlds:
        cnt = np.asarray(w.download(abi.PART_COUNTERS)).view(np.uint32).copy()
        words = cnt[:11 * W * H].reshape(H, W, 11)
        surface = H - 1 - 28
        moves = [((30, surface + 3), (51, surface + 6)), ((31, surface + 3), (52, surface + 7)),
                 ((10, surface + 10), (12, surface + 10)), ((40, surface + 12), (41, surface + 13))]
        for (x, y), (tx, ty) in moves:
            words[y, x, 4] = 0x80000000 | (ty * W + tx)
        w.upload(abi.PART_COUNTERS, cnt)


@pytest.mark.parametrize("mode", ["reference", "relative", "viscous"])
def test_bodies_match_oracle(mode):
    cells, motion, swept, kinds, props = scenes.bodies_tank(W, H)
    if mode == "viscous":
        props[E.WATER, 0] = 0.05          # a viscous water: the sweeps, the bodies' drag and faces
    solver = SolverSettings.reference() if mode == "reference" else SolverSettings.relative(1e-6)
    wo, wg = world("oracle", cells, motion, swept, props, kinds), world("gpu", cells, motion, swept, props, kinds)
    so, sg = make_sim(wo, solver), make_sim(wg, solver)
    if mode == "viscous":
        so.fluid_stats["viscous"] = sg.fluid_stats["viscous"] = 1
    for s in range(6):
        say(f"[bodies {mode}] step {s}")
        if s == 3:
            push_aside((wo, wg), so.frame_no)
        so.step()
        sg.step()
        assert np.array_equal(so.last_stats, sg.last_stats), (s, so.last_stats, sg.last_stats)
        d = body_diffs(wo, wg)
        assert not d, f"step {s}: " + "; ".join(d)
    # the liquid did something to every body
    m = np.asarray(wg.download(abi.PART_MOTION)).view(np.int32)
    for slot in (3, 7, 12, 20):
        assert np.any(m[abi.liquid_slot(slot, 0):abi.liquid_slot(slot, 3)] != 0), slot
    wo.close()
    wg.close()


def test_bodies_in_a_graphed_batch():
    """Two bodies tanks and a plain dam stepped as one batch with graphs,
    each against the same world stepped alone with plain launches."""
    from worldinagrain_amd.gpu import GpuBatch, GpuWorld
    cells, motion, swept, kinds, props = scenes.bodies_tank(W, H)
    plain = scenes.dam_break(W, H, block=(30, 30))
    def pair(c, bodies):
        out = []
        for _ in range(2):
            w = GpuWorld(W, H, kinds=kinds)
            w.set_element_props(props)
            w.upload(abi.PART_CELLS, c)
            if bodies:
                w.upload(abi.PART_MOTION, motion)
                w.upload(abi.PART_SWEPT, swept)
            s = Simulation(w, W, H, solver=SolverSettings.relative(1e-6))
            s.bodies = s.bodies_run = bodies
            out.append((w, s))
        return out
    specs = [pair(cells, True), pair(cells, True), pair(plain, False)]
    specs[1][0][1].seed(99)
    specs[1][1][1].seed(99)
    batch = GpuBatch([p[0][0] for p in specs])
    batch.set_graphs(True)
    for s in range(4):
        say(f"[bodies batch] step {s}")
        step_batch(batch, [p[0][1] for p in specs])
        for p in specs:
            p[1][1].step()
        for k, p in enumerate(specs):
            d = body_diffs(p[1][0], p[0][0])
            assert not d, f"world {k} step {s}: " + "; ".join(d)
    batch.close()
