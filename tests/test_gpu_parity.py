"""The HIP liquid step against the oracle (the CPU restatement of
fluid.wgsl), on the same seeded worlds, step by step.  Both compute the CG
sums in FP64 in the same tree order, so every cell, face and counter word
must be bit-identical, and the particles identical as multisets (their slot
order in `parts` is whatever order the GPU's threads compacted them in, as in
the reference)."""
import sys

import numpy as np
import pytest

from worldinagrain_amd import abi, scenes
from worldinagrain_amd import elements as E
from worldinagrain_amd.host import Simulation
from tests.helpers import make_oracle, make_gpu, compare_worlds, config_sim

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def run_pair(name, W, H, cells, steps, mode="reference", check_every=1, coarse=0, tol=1e-6):
    say(f"[parity] {name} {W}x{H} {mode} steps={steps}")
    wo = make_oracle(W, H, cells, coarse)
    wg = make_gpu(W, H, cells, coarse)
    if mode == "reference":
        so, sg = Simulation(wo, W, H), Simulation(wg, W, H)
    else:
        so, sg = config_sim(wo, W, H, tol=tol), config_sim(wg, W, H, tol=tol)
    for s in range(steps):
        so.step()
        sg.step()
        assert np.array_equal(so.last_stats, sg.last_stats), (s, so.last_stats, sg.last_stats)
        if (s + 1) % check_every == 0 or s == steps - 1:
            diffs = compare_worlds(wo, wg)
            assert not diffs, f"step {s}: " + "; ".join(diffs)
    return wo, wg


def test_first_step_dam():
    W = H = 64
    run_pair("first step", W, H, scenes.dam_break(W, H, block=(32, 40)), 1)


def test_dam_reference_settings_40_steps():
    W, H = 64, 48
    run_pair("dam", W, H, scenes.dam_break(W, H, block=(30, 36)), 40, check_every=10)


def test_odd_world_size():
    W, H = 61, 37
    run_pair("odd", W, H, scenes.dam_break(W, H, block=(29, 30)), 12, check_every=4)


def test_discs_config_mode():
    W = H = 128
    cells, _ = scenes.discs_dam(W, H, discs=20, rmin=3, rmax=10)
    run_pair("discs", W, H, cells, 6, mode="config", check_every=3)


def test_tiled_levels_odd_size():
    """330 x 250: coarse levels 1 and 2 (165 x 125, 83 x 63) are swept in
    tiles with partial tiles at their edges, 3 and 4 in one workgroup."""
    W, H = 330, 250
    cells, _ = scenes.discs_dam(W, H, discs=40, rmin=3, rmax=12)
    wo, wg = run_pair("tiled levels", W, H, cells, 4, mode="config", check_every=2)
    with pytest.raises(Exception):
        wg.set_tile(3, 16)          # level 3 is in the one-workgroup kernel


def test_tiled_levels_reference_settings():
    W, H = 300, 260
    run_pair("tiled levels, reference", W, H, scenes.dam_break(W, H, block=(140, 150)), 6, check_every=3)


def test_tile_sides_do_not_change_results():
    """The tiles' halos make their sweeps the whole-level passes': any tile
    side, any level for the one-workgroup kernel to start at, and any split
    of its levels between the workgroup and one wavefront gives the same
    bits; so does a clone of a world stepped alike."""
    W, H = 330, 250
    cells, _ = scenes.discs_dam(W, H, discs=40, rmin=3, rmax=12)
    wa, wb, wl = make_gpu(W, H, cells), make_gpu(W, H, cells), make_gpu(W, H, cells)
    sa, sb, sl = config_sim(wa, W, H), config_sim(wb, W, H), config_sim(wl, W, H)
    wb.set_tile(1, 8)
    wb.set_tile(2, 32)
    wb.set_coarse_wave(0)                    # every coarse level by the whole workgroup
    wa.set_coarse_wave(1 << 20)              # every one below the first by one wavefront
    wl.set_lds_level(wl.layout.levels)       # every level but the coarsest in tiles
    wl.set_tile(3, 8)
    with pytest.raises(Exception):
        wl.set_lds_level(1)                  # levels 1 .. 4 do not fit one workgroup's LDS
    for s in range(3):
        sa.step()
        sb.step()
        sl.step()
    diffs = compare_worlds(wa, wb) + compare_worlds(wa, wl)
    assert not diffs, "; ".join(diffs)
    wc = wa.clone()
    import copy
    sc = copy.copy(sa)
    sc.world = wc
    sa.step()
    sc.step()
    diffs = compare_worlds(wa, wc)
    assert not diffs, "; ".join(diffs)


def test_still_tank_gpu():
    W, H = 32, 32
    wo, wg = run_pair("still tank", W, H, scenes.still_tank(W, H, 14), 30, check_every=10)
    f = wg.download(abi.PART_FIELD)[:8 * W * H].reshape(H, W, 8)
    col = f[:, W // 2, 0][f[:, W // 2, 2] > 0]
    assert np.all(np.abs(np.diff(col) - float(E.GRAVITY)) < 0.05 * float(E.GRAVITY))


@pytest.mark.parametrize("W,H,block,mode", [
    (8, 8, (5, 6), "reference"),        # one block, one level
    (5, 3, (4, 3), "reference"),        # smaller than a block
    (200, 9, (120, 8), "config"),       # a strip one block and a row high
    (9, 130, (8, 100), "config"),       # a column
])
def test_small_and_thin_worlds(W, H, block, mode):
    run_pair(f"small {W}x{H}", W, H, scenes.dam_break(W, H, block=block), 6, mode=mode, check_every=2)


def test_full_size_discs_world():
    # the benchmark's world at its full size (1024 x 1024, 200 discs, 1.47 M
    # particles, MG-PCG to 1e-6): the oracle takes about 1.5 s a step
    W = H = 1024
    cells, _ = scenes.discs_dam(W, H)
    run_pair("full size discs", W, H, cells, 2, mode="config", check_every=1)


def test_idle_world():
    W = H = 32
    cells = scenes.empty_cells(W, H)
    wo, wg = run_pair("idle", W, H, cells, 3)
    assert int(wg.report()[abi.STAT_ACTIVE]) == 0


def test_particle_overflow():
    """Drawn cells that count 255 particles with no home get 255 seeded each
    (syncSeed, fluid.wgsl:876-877): past capacity, the count says how many
    tried and the rest are lost (fluid.wgsl:105-107)."""
    W = H = 16
    cells = scenes.empty_cells(W, H)
    m = np.zeros((H, W), bool)
    m[4:8, 4:8] = True
    scenes.paint(cells, W, H, m, E.WATER)
    words = cells[:4 * W * H].reshape(H, W, 4)
    words[m, 0] |= np.uint32(E.FL_DRAWN << 28)
    words[m, 3] = (np.uint32(0xffffff) << 8) | np.uint32(255)
    say("[parity] overflow")
    wo, wg = make_oracle(W, H, cells), make_gpu(W, H, cells)
    so, sg = Simulation(wo, W, H), Simulation(wg, W, H)
    so.step()
    sg.step()
    cap = wg.layout.capacity
    assert int(wg.download(abi.PART_PARTICLES1)[0]) == 16 * 255 > cap
    assert int(wo.part(abi.PART_PARTICLES1)[0]) == 16 * 255
    assert int(wg.report()[abi.STAT_PARTICLES]) == int(wo.report()[abi.STAT_PARTICLES]) <= cap


def test_pressure_against_fp64_cg():
    W = H = 96
    cells = scenes.dam_break(W, H, block=(48, 70))
    say("[parity] pressure vs fp64")
    wo, wg = make_oracle(W, H, cells), make_gpu(W, H, cells)
    so, sg = config_sim(wo, W, H, tol=1e-6), config_sim(wg, W, H, tol=1e-6)
    for _ in range(4):
        so.step()
        sg.step()
    wo.capture(True)
    so.step()
    sg.step()
    b = wo.captured_b()
    x64, it, rel = wo.cg64(b, tol=1e-12)
    f = wg.download(abi.PART_FIELD)[:8 * W * H].reshape(-1, 8)
    fl = f[:, 2] > 0
    err = np.linalg.norm(f[fl, 0] - x64[fl]) / np.linalg.norm(x64[fl])
    say(f"  relative error vs FP64 CG: {err:.3e}")
    assert err < 1e-5
    s = wg.solve_stats()
    assert s.rr <= 1e-12 * s.bb
