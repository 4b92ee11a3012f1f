"""The air's projection on the GPU (wig_air_project) against the oracle
(oracle/air_oracle.c): every buffer bit-identical after each solve (the
levels, the CG's cells and sums, the status slots), over grid shapes the
level rule treats differently (one level only, thin, odd sides, a single
cell), both LDS tile sides, several warm-started steps, the zones' mean,
iterations cut short, and the full 1024 x 1024 grid."""
import sys

import numpy as np
import pytest

from oracle.air import OracleAir
from worldinagrain_amd import scenes
from worldinagrain_amd.air import GpuAir, AirProjection, PART_MG, PART_CG, PART_STATUS

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def same(g, o, what):
    mg, cg, st = g.download(PART_MG), g.download(PART_CG), g.download(PART_STATUS)
    for name, a, b in (("mg", mg, o.mg), ("cg", cg, o.cg), ("status", st, o.status)):
        d = np.nonzero(a.view(np.uint32) != b.view(np.uint32))[0]
        assert d.size == 0, f"{what}: {name} differs at {d.size} words, first {d[:6]}: {a[d[:3]]} vs {b[d[:3]]}"


def next_step(g, o, W, H, seed, step):
    """gather's and divergence's level 0 for another step's flow, the rest
    (coarse levels, the warm potentials) as the last solve left them."""
    mg0, _ = scenes.air_rooms(W, H, seed=seed, flow_seed=1000 + step)
    n = W * H
    mg = g.download(PART_MG)
    mg[:8 * n] = mg0.reshape(-1)
    g.upload(PART_MG, mg)
    o.mg[:8 * n] = mg0.reshape(-1)


@pytest.mark.parametrize("W,H,seed,tile,coarse", [
    (96, 80, 3, 16, 1), (96, 80, 3, 32, 1), (96, 80, 3, 0, 1), (96, 80, 3, 16, 0),
    (200, 130, 5, 16, 1), (200, 130, 5, 32, 1), (200, 130, 5, 0, 1), (200, 130, 5, 0, 0),
    (48, 40, 6, 0, 1), (37, 29, 4, 32, 1), (16, 12, 2, 16, 1), (5, 190, 4, 16, 1), (5, 190, 4, 16, 0),
    (1, 1, 0, 16, 1), (300, 7, 8, 32, 1), (130, 70, 9, 0, 1), (96, 80, 3, 8, 0), (200, 130, 5, 8, 0),
])
def test_air_projection_matches_oracle(W, H, seed, tile, coarse):
    # coarse: the levels that fit 64 x 64 in one kernel (1, opt-in) or
    # a launch each (0)
    say(f"[air] {W}x{H} seed {seed} tile {tile} coarse {coarse}")
    mg0, x0 = scenes.air_rooms(W, H, seed=seed)
    g, o = GpuAir(W, H), OracleAir(W, H)
    assert g.dims == o.dims
    g.set_coarse(coarse)
    if tile:                      # (0: the default, 32 on level 0, 16 on level 1, 8 below)
        g.set_tile(tile)
    g.load(mg0, x0)
    o.load(mg0, x0)
    for step in range(3):
        g.project(64)
        o.project(64)
        g.sync()
        same(g, o, f"step {step}")
        say(f"[air]   step {step}: {int(o.status[4])} iterations")
        next_step(g, o, W, H, seed, step)
    g.close()


def test_air_mean_and_cut_short():
    W, H = 120, 90
    say(f"[air] {W}x{H} mean, cycles 3")
    rng = np.random.default_rng(2)
    mg0, x0 = scenes.air_rooms(W, H, seed=7)
    mean = np.where(mg0[..., 7] > 0, rng.normal(0, 1e-4, (H, W)), 0.0).astype(np.float32)
    g, o = GpuAir(W, H), OracleAir(W, H)
    g.load(mg0, x0, mean)
    o.load(mg0, x0, mean)
    g.project(3)
    o.project(3)
    g.sync()
    same(g, o, "3 iterations")
    assert o.status[4] == 3
    g.close()


def test_air_host_policy_on_gpu():
    # the host's airCycles over steps: the same iterations asked and run
    W, H = 128, 96
    say(f"[air] {W}x{H} host policy")
    mg0, x0 = scenes.air_rooms(W, H, seed=12)
    g, o = GpuAir(W, H), OracleAir(W, H)
    g.load(mg0, x0)
    o.load(mg0, x0)
    pg, po = AirProjection(g), AirProjection(o)
    for step in range(4):
        assert pg.cycles == po.cycles
        assert pg.step() == po.step()
        same(g, o, f"step {step}")
        next_step(g, o, W, H, 12, step)
    g.close()


def test_air_full_size():
    W = H = 1024
    say(f"[air] {W}x{H} default tiles")
    mg0, x0 = scenes.air_rooms(W, H, seed=21, walls=60)
    g, o = GpuAir(W, H), OracleAir(W, H)
    g.load(mg0, x0)
    o.load(mg0, x0)
    g.project(64)
    o.project(64)
    g.sync()
    same(g, o, "1024x1024")
    its = int(o.status[4])
    say(f"[air]   {its} iterations")
    assert 0 < its < 64
    r = g.download(PART_CG)[:4 * W * H].reshape(H, W, 4)[..., 1]
    assert np.abs(r).max() < 1e-4
    g.close()


@pytest.mark.parametrize("W,H,tile", [(96, 80, 0), (150, 110, 16)])
def test_air_fractional_weights(W, H, tile):
    # faces of any weight (gather's are 0 or 1): the same bits as the oracle
    say(f"[air] {W}x{H} fractional weights, tile {tile}")
    mg0, x0 = scenes.air_rooms(W, H, seed=17, fractional=True)
    g, o = GpuAir(W, H), OracleAir(W, H)
    if tile:
        g.set_tile(tile)
    g.load(mg0, x0)
    o.load(mg0, x0)
    g.project(64)
    o.project(64)
    g.sync()
    same(g, o, "fractional")
    assert int(o.status[4]) > 0
    g.close()
