"""The air projection's oracle (oracle/air_oracle.c), pinned by what the
mathematics fixes, as the reference holds no tests or vectors for it:

- the level rule and buffer sizes (js/sim.js:712-723,858-864) in closed form;
- the coarse operators are the Galerkin products of the finer ones over the
  piecewise-constant prolongation (restrictCell, shaders/air.wgsl:1341-1417):
  level 1's P^T A0 P / COARSEN, the deeper ones P^T A P, assembled here
  independently with scipy from the levels' face weights;
- the restricted right-hand side is P^T of the smoothed level's residual;
- the V-cycle (the preconditioner) is linear, symmetric and positive;
- the first iteration's sums are r.z and r.z / p.Ap in float64 (to float32's
  rounding), beta 0;
- the solve ends with its recurrence residual the true one of A x = -b, under
  the tolerance;
- an iteration after convergence changes no buffer;
- the host's iteration policy (js/sim.js:1916-1920)."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy import ndimage

from oracle.air import OracleAir
from worldinagrain_amd import scenes
from worldinagrain_amd.air import AirProjection, AIR_CYCLES

ANCHOR, COARSEN = 1e-6, 1.8


def test_levels_and_sizes():
    # halve, rounding up, until <= 256 cells (js/sim.js:712-719)
    assert OracleAir(1024, 1024).dims == [(1024, 1024), (512, 512), (256, 256), (128, 128), (64, 64), (32, 32), (16, 16)]
    assert OracleAir(96, 80).dims == [(96, 80), (48, 40), (24, 20), (12, 10)]
    assert OracleAir(16, 16).dims == [(16, 16)]
    assert OracleAir(5, 190).dims == [(5, 190), (3, 95), (2, 48)]
    o = OracleAir(200, 130)
    assert o.dims == [(200, 130), (100, 65), (50, 33), (25, 17), (13, 9)]
    cells = sum(w * h for w, h in o.dims)
    assert o.mg.size == 8 * cells
    groups = 25 * 17
    assert o.cg.size == 4 * (200 * 130 + -(-(4 + groups) // 4))


def anchor_at(l, x, y, W, H):
    sx = min((x + 1) << l, W) - (x << l)
    sy = min((y + 1) << l, H) - (y << l)
    return ANCHOR * sx * sy / (COARSEN if l >= 1 else 1.0)


def assemble(level, l, W, H):
    """A level's operator on its free (not held) cells, from its face weights:
    faces between free cells couple them, every other face (to a held cell,
    the world's edges) and the anchor is on the diagonal."""
    h, w, _ = level.shape
    held = level[..., 5] > 0
    idx = -np.ones((h, w), int)
    free = np.argwhere(~held)
    idx[~held] = np.arange(len(free))
    A = sp.lil_matrix((len(free), len(free)))
    for y, x in free:
        i = idx[y, x]
        d = float(level[y, x, 2]) + float(level[y, x, 3]) + float(level[y, x, 4]) + anchor_at(l, x, y, W, H)
        if x > 0:
            d += float(level[y, x - 1, 2])
        if y > 0:
            d += float(level[y - 1, x, 3])
        A[i, i] = d
        for (nx, ny, wgt) in ((x + 1, y, level[y, x, 2]), (x, y + 1, level[y, x, 3]),
                              (x - 1, y, level[y, x - 1, 2] if x > 0 else 0), (x, y - 1, level[y - 1, x, 3] if y > 0 else 0)):
            if 0 <= nx < w and 0 <= ny < h and idx[ny, nx] >= 0:
                A[i, idx[ny, nx]] = -float(wgt)
    return A.tocsr(), idx


def prolongation(idx_f, idx_c):
    """Piecewise constant: a coarse cell's correction on its free fine cells."""
    hf, wf = idx_f.shape
    P = sp.lil_matrix((int(idx_f.max()) + 1, int(idx_c.max()) + 1))
    for y in range(hf):
        for x in range(wf):
            if idx_f[y, x] >= 0 and idx_c[y // 2, x // 2] >= 0:
                P[idx_f[y, x], idx_c[y // 2, x // 2]] = 1.0
    return P.tocsr()


def started(W, H, seed, mean=None):
    mg0, x0 = scenes.air_rooms(W, H, seed=seed)
    o = OracleAir(W, H)
    o.load(mg0, x0, mean)
    o.start()
    return o, mg0, x0


@pytest.mark.parametrize("W,H,seed,fractional", [(40, 36, 1, False), (37, 29, 4, False), (41, 33, 8, True)])
def test_coarse_operators_are_galerkin(W, H, seed, fractional):
    mg0, x0 = scenes.air_rooms(W, H, seed=seed, fractional=fractional)
    o = OracleAir(W, H)
    o.load(mg0, x0)
    o.start()
    # pin a cell so a coarse cell has a held one among its four
    o.level(0)[3, 4, 5] = 1.0
    o.vcycle()
    ops = [assemble(o.level(l), l, W, H) for l in range(o.levels)]
    for l in range(o.levels - 1):
        (Af, idf), (Ac, idc) = ops[l], ops[l + 1]
        P = prolongation(idf, idc)
        G = (P.T @ Af @ P).toarray() / (COARSEN if l == 0 else 1.0)
        assert np.abs(G - Ac.toarray()).max() < 2e-6 * max(1.0, np.abs(G).max()), l


@pytest.mark.parametrize("W,H,seed", [(40, 36, 2), (33, 47, 5)])
def test_restricted_rhs_is_the_residual(W, H, seed):
    o, _, _ = started(W, H, seed)
    before = o.level(0)[..., 1].copy()          # level 0's rhs (-r) as the V-cycle takes it
    o.vcycle()
    pre = o.cells()[..., 3]                     # the smoothed correction down left (cg's w)
    lv0 = o.level(0).copy()
    lv0[..., 1] = before
    A0, id0 = assemble(lv0, 0, W, H)
    _, id1 = assemble(o.level(1), 1, W, H)
    P = prolongation(id0, id1)
    free = id0 >= 0
    res = before[free].astype(np.float64) + A0 @ pre[free].astype(np.float64)
    want = P.T @ res
    got = o.level(1)[..., 1][id1 >= 0]
    assert np.abs(got - want).max() < 1e-5 * max(1.0, np.abs(want).max())
    # the smoothing took most of the residual off
    assert np.linalg.norm(res) < 0.8 * np.linalg.norm(before[free])


def vcycle_of(o, state, v):
    """z = M v: the V-cycle on rhs -v from a clean state (corrections none)."""
    o.mg[:] = state[0]
    o.cg[:] = state[1]
    lv = o.level(0)
    lv[..., 0] = 0.0
    lv[..., 1] = -v
    o.cells()[..., 3] = 0.0
    o.vcycle()
    return o.level(0)[..., 0].astype(np.float64).copy()


@pytest.mark.parametrize("W,H", [(48, 40), (16, 12)])
def test_vcycle_symmetric_positive(W, H):
    o, _, _ = started(W, H, 3)
    state = (o.mg.copy(), o.cg.copy())
    free = ~(o.level(0)[..., 5] > 0)
    rng = np.random.default_rng(0)
    u = (rng.normal(size=(H, W)) * free).astype(np.float32)
    v = (rng.normal(size=(H, W)) * free).astype(np.float32)
    Mu, Mv = vcycle_of(o, state, u), vcycle_of(o, state, v)
    a, b = float((u * Mv).sum()), float((v * Mu).sum())
    assert abs(a - b) < 1e-4 * max(abs(a), abs(b))
    assert float((u * Mu).sum()) > 0 and float((v * Mv).sum()) > 0
    # linear: M(2u + v) = 2 M u + M v
    Mw = vcycle_of(o, state, (2 * u + v).astype(np.float32))
    assert np.abs(Mw - (2 * Mu + Mv)).max() < 1e-4 * np.abs(Mw).max()


def test_first_iteration_sums():
    W, H = 56, 44
    o, _, _ = started(W, H, 6)
    r0 = o.cells()[..., 1].astype(np.float64).copy()
    snap = (o.mg.copy(), o.cg.copy(), o.status.copy())
    o.vcycle()
    z0 = o.level(0)[..., 0].astype(np.float64).copy()
    o.mg[:], o.cg[:], o.status[:] = snap
    o.iteration(0)
    s = o.sums()
    rho = float((r0 * z0).sum())
    assert s[2] == 0.0                                   # beta, the first iteration
    assert abs(s[0] - rho) < 1e-5 * abs(rho)
    A0, id0 = assemble(o.level(0), 0, W, H)
    free = id0 >= 0
    p = np.where(free, z0, 0.0)
    Ap = np.zeros_like(p)
    Ap[free] = A0 @ p[free]
    assert np.abs(o.cells()[..., 3] - Ap).max() < 1e-5 * np.abs(Ap).max()
    assert abs(s[1] - rho / float((p * Ap).sum())) < 1e-4 * abs(s[1])
    # each 8 x 8 group's partial is its share of p . Ap
    gx = -(-W // 8)
    part = (p * Ap)
    for g in (0, 5, gx + 2):
        by, bx = divmod(g, gx)
        assert abs(s[4 + g] - part[by * 8:by * 8 + 8, bx * 8:bx * 8 + 8].sum()) < 1e-5 * max(1e-3, abs(part).sum() / 10)


@pytest.mark.parametrize("W,H,seed,with_mean", [(96, 80, 3, False), (64, 64, 9, True), (16, 12, 2, False)])
def test_solve_meets_its_tolerance(W, H, seed, with_mean):
    rng = np.random.default_rng(seed)
    mg0, x0 = scenes.air_rooms(W, H, seed=seed)
    mean = None
    if with_mean:
        # a source in every cell with gas, and (meanExpansion, air.wgsl:
        # 1006-1021) each sealed room's mean of it taken off there: a room
        # held by no edge and no black hole sums to none
        carries, held = mg0[..., 7] > 0, mg0[..., 5] > 0
        src = np.where(carries & ~held, rng.normal(0, 1e-3, (H, W)), 0.0).astype(np.float32)
        mg0[..., 1] += src
        rooms, n = ndimage.label(carries & ~held)
        mean = np.zeros((H, W), np.float32)
        near_held = ndimage.binary_dilation(held) & ~held
        for k in range(1, n + 1):
            room = rooms == k
            opened = (mg0[..., 4][room] > 0).any() or room[:, -1].any() or room[-1, :].any() or (room & near_held).any()
            if not opened:
                mean[room] = -src[room].astype(np.float64).mean()
    o = OracleAir(W, H)
    o.load(mg0, x0, mean)
    o.project(64)
    its = int(o.status[4])
    assert 0 < its < 64
    # the true residual of A x = -b (b with the zones' mean where gas), held
    # cells at their potential
    b = mg0[..., 1].astype(np.float64)
    if mean is not None:
        b = b + np.where(mg0[..., 7] > 0, mean, 0.0)
    x = o.cells()[..., 0].astype(np.float64)
    right, down, edge = (mg0[..., k].astype(np.float64) for k in (2, 3, 4))
    left = np.zeros_like(right)
    left[:, 1:] = right[:, :-1]
    up = np.zeros_like(down)
    up[1:, :] = down[:-1, :]
    diag = right + down + edge + ANCHOR + left + up
    nb = np.zeros_like(x)
    nb[:, :-1] += right[:, :-1] * x[:, 1:]
    nb[:-1, :] += down[:-1, :] * x[1:, :]
    nb[:, 1:] += left[:, 1:] * x[:, :-1]
    nb[1:, :] += up[1:, :] * x[:-1, :]
    held = mg0[..., 5] > 0
    r_true = np.where(held, 0.0, -b - (diag * x - nb))
    r = o.cells()[..., 1]
    assert np.abs(r_true - r).max() < 2e-5
    assert np.abs(r).max() < 1e-4
    assert np.abs(r_true).max() < 1.2e-4


def test_converged_iterations_change_nothing():
    W, H = 72, 60
    mg0, x0 = scenes.air_rooms(W, H, seed=11)
    o = OracleAir(W, H)
    o.load(mg0, x0)
    o.project(64)
    its = int(o.status[4])
    a = OracleAir(W, H)
    a.load(mg0, x0)
    a.project(its + 1)
    b = OracleAir(W, H)
    b.load(mg0, x0)
    b.project(its + 6)
    assert np.array_equal(a.mg.view(np.uint32), b.mg.view(np.uint32))
    assert np.array_equal(a.cg.view(np.uint32), b.cg.view(np.uint32))
    assert a.status[4] == b.status[4] == its
    # cut short, the iterations run are all that were given
    c = OracleAir(W, H)
    c.load(mg0, x0)
    c.project(3)
    assert c.status[4] == 3 and np.abs(c.cells()[..., 1]).max() > 1e-4


def test_cycles_policy():
    class Fake:
        def __init__(self, runs):
            self.runs = list(runs)
            self.status = np.zeros(8, np.uint32)
            self.given = []

        def project(self, cycles):
            self.given.append(cycles)
            self.status[4] = min(self.runs.pop(0), cycles)

    f = Fake([20, 10, 40, 3])
    p = AirProjection(f)
    for _ in range(4):
        p.step()
    # 64 first; then 2 x 20 + 8; 2 x 10 + 8; 28 used up: 64 again
    assert f.given == [AIR_CYCLES, 48, 28, AIR_CYCLES]
    assert p.cycles == 2 * 3 + 8
