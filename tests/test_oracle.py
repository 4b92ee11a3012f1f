"""The oracle (the CPU restatement of fluid.wgsl) pinned without the
reference running: its integer functions against independent restatements
and library routines, its float helpers against closed forms, and whole steps
against the reference's own scene specifications (js/scenes.js:398-407,
1084-1091: the dam loses not a drop; the still tank comes to rest with its
pressure growing with depth)."""
import numpy as np
import pytest

from worldinagrain_amd import abi, scenes
from worldinagrain_amd import elements as E
from worldinagrain_amd.host import Simulation, SolverSettings, mulberry32_next
from tests.helpers import make_oracle, config_sim


def test_f16_matches_numpy(oracle_lib):
    """pack2x16float's conversion, round to nearest even, against numpy."""
    rng = np.random.default_rng(1)
    xs = np.concatenate([
        rng.normal(0, 1, 20000), rng.normal(0, 1e-5, 5000), rng.normal(0, 3e4, 5000),
        [0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e9, -1e9, 2.0 ** -24, 2.0 ** -25, 1.5 * 2.0 ** -25, 2.0 ** -14,
         np.inf, -np.inf],
    ]).astype(np.float32)
    for x in xs:
        with np.errstate(over="ignore"):
            want = int(np.float16(x).view(np.uint16))
        assert oracle_lib.wor_f16(float(x)) == want, x
    for h in range(0, 65536, 7):
        v = np.uint16(h).view(np.float16)
        if np.isnan(v):
            continue
        assert np.float32(oracle_lib.wor_f32(h)) == np.float32(v), h


def _pcg_numpy(v):
    """The PCG RXS-M-XS hash (Jarzynski & Olano 2020), in numpy's u32."""
    v = np.uint32(v)
    with np.errstate(over="ignore"):
        s = v * np.uint32(747796405) + np.uint32(2891336453)
        w = ((s >> ((s >> np.uint32(28)) + np.uint32(4))) ^ s) * np.uint32(277803737)
    return int((w >> np.uint32(22)) ^ w)


def test_pcg_known_answers(oracle_lib):
    for v in [0, 1, 2, 12345, 0xdeadbeef, 0xffffffff, 0x5eed]:
        assert oracle_lib.wor_pcg(v) == _pcg_numpy(v)
    # a known answer computed once from the published definition
    assert _pcg_numpy(0) == 129708002


def test_seed_positions_quarters(oracle_lib):
    """seedPos (fluid.wgsl:845-855): four seeds in the cell's quarters, each
    0.3 * U[-0.5, 0.5) from its quarter's middle."""
    import ctypes as ct
    out = (ct.c_float * 8)()
    for (x, y, i) in [(5, 7, 1000), (0, 0, 0), (100, 3, 77)]:
        oracle_lib.wor_seed_pos(x, y, i, 1, 0x1234, 4, out)
        p = np.array(out).reshape(4, 2)
        quarters = np.array([[0, 0], [1, 1], [1, 0], [0, 1]])
        mid = np.array([x, y]) + 0.25 + 0.5 * quarters
        assert np.all(np.abs(p - mid) <= 0.15 + 1e-6)
        assert np.all((p >= [x, y]) & (p < [x + 1, y + 1]))


def test_exact_sums_round_trip(oracle_lib):
    import ctypes as ct
    out = (ct.c_uint32 * 2)()
    rng = np.random.default_rng(3)
    for x in rng.normal(0, 100, 200).astype(np.float32):
        oracle_lib.wor_exact_parts(float(x), 4294967296.0, out)
        v = (int(out[1]) << 32 | int(out[0]))
        if v >= 1 << 63:
            v -= 1 << 64
        assert v == int(np.round(np.float64(x) * 2.0 ** 32))   # x * 2^32 is exact in f32: round is the identity
        assert np.float32(oracle_lib.wor_exact_value(out[0], out[1], 4294967296.0)) == x


def test_fine_beta_closed_forms(oracle_lib):
    fb = oracle_lib.wor_fine_beta
    assert fb(1.0, 1.0, 1.0, 1.0) == 1.0                  # 2 / (a + b)
    assert fb(1.0, 0.0, 0.0, 0.0) == 2.0                  # liquid with no particles: half a cell to the surface
    assert np.isclose(fb(1.0, 0.0, 1.0, 0.0), 1.0 / 0.5)  # theta = (1 - .5) / 1
    assert np.isclose(fb(1.0, 0.0, 0.52, 0.3), 10.0)      # theta 0.02 / 0.22 clamped at THETA_MIN 0.1
    assert np.isclose(fb(1.0, 0.0, 0.8, 0.2), 1.0 / 0.5)   # theta 0.3 / 0.6
    assert fb(1.0, -1.0, 1.0, 1.0) == 0.0 and fb(0.0, 0.0, 1.0, 1.0) == 0.0
    assert fb(2.0, 0.0, 1.0, 0.0) == fb(0.0, 2.0, 0.0, 1.0)   # symmetric: A is


def test_mulberry32_against_javascript_semantics():
    """The host's mulberry32 (js/sim.js:1390-1397) against a restatement that
    follows JavaScript's Math.imul and >>> on signed 32-bit ints."""
    def imul(a, b):
        r = (a * b) & 0xFFFFFFFF
        return r - (1 << 32) if r >= 1 << 31 else r

    def js_random(state):
        state = (state + 0x6d2b79f5) & 0xFFFFFFFF
        t = state
        t = imul(t ^ (t >> 15), t | 1) & 0xFFFFFFFF
        t = (t ^ ((t + (imul(t ^ (t >> 7), t | 61) & 0xFFFFFFFF)) & 0xFFFFFFFF)) & 0xFFFFFFFF
        return state, (t ^ (t >> 14)) & 0xFFFFFFFF

    a = b = 0x5EED
    for _ in range(100):
        a, ra = mulberry32_next(a)
        b, rb = js_random(b)
        assert ra == rb


def test_layout_sizes():
    from oracle.oracle import OracleWorld
    w = OracleWorld(256, 256)
    L = w.layout
    assert L.levels == 4 and (L.level_width[4], L.level_height[4]) == (16, 16)       # 5 levels in all
    C, F = 256 * 256, 257 * 256 * 2
    assert L.faces == F and L.capacity == 5 * C
    levels = 128 ** 2 + 64 ** 2 + 32 ** 2 + 16 ** 2
    assert L.part_bytes[abi.PART_FIELD] == (C + levels) * 32 + (F + 2 * ((F + 3) // 4)) * 16   # js/sim.js:772-773
    assert L.part_bytes[abi.PART_COUNTERS] == (11 * C + 2 * F) * 4
    assert L.part_bytes[abi.PART_PARTICLES0] == 16 + 48 * 5 * C
    assert OracleWorld(1024, 1024).layout.levels == 6                                   # 7 in all
    assert OracleWorld(61, 37).layout.blocks == 8 * 5


def test_splat_weights_sum_to_a_cell():
    """One particle's splat: per axis the weight words sum to FIX within the
    four roundings, and its momentum words to vel * FIX (no gradient)."""
    W = H = 16
    cells = scenes.empty_cells(W, H)
    m = np.zeros((H, W), bool)
    m[8, 8] = True
    scenes.paint(cells, W, H, m, E.WATER, vel=(np.full((H, W), 0.25), np.full((H, W), -0.125)))
    w = make_oracle(W, H, cells)
    sim = Simulation(w, W, H, earth_g=0.0)
    sim.step()
    # the second step's gather reads what the first seeded
    cnt = w.part(abi.PART_COUNTERS)
    C = W * H
    F = (W + 1) * H + W * (H + 1)
    sums = cnt[11 * C:11 * C + 2 * F].reshape(F, 2).view(np.int32)
    nu = (W + 1) * H
    for lo, hi, vel in ((0, nu, 0.25), (nu, F, -0.125)):
        wsum = int(sums[lo:hi, 1].sum())
        assert abs(wsum - 4 * 65536) <= 4 * 4   # 4 particles, 4 faces each
        msum = int(sums[lo:hi, 0].sum())
        assert abs(msum - vel * 4 * 65536) <= 16


def test_dam_break_conserves_particles():
    """Dam (js/scenes.js:1088-1091): not a drop is lost."""
    W, H = 48, 40
    w = make_oracle(W, H, scenes.dam_break(W, H, block=(20, 24)))
    sim = Simulation(w, W, H)
    counts = []
    for _ in range(60):
        sim.step()
        counts.append(int(w.part(abi.PART_PARTICLES0)[0]))
    assert counts[0] == (20 - 1) * (24 - 1) * 4
    assert all(c == counts[0] for c in counts)
    f = w.part(abi.PART_FIELD)
    assert np.isfinite(f).all()


def test_still_tank_hydrostatic():
    """Still tank (js/scenes.js:398-402,1084-1087): comes to rest, its
    pressure growing with depth at rho g a cell (columnWeight's)."""
    W, H, depth = 24, 24, 12
    w = make_oracle(W, H, scenes.still_tank(W, H, depth))
    sim = Simulation(w, W, H)
    for _ in range(40):
        sim.step()
    f = w.part(abi.PART_FIELD)[:8 * W * H].reshape(H, W, 8)
    g = float(E.GRAVITY)
    col = f[:, W // 2, 0]
    rho = f[:, W // 2, 2]
    fluid_rows = np.nonzero(rho > 0)[0]
    p = col[fluid_rows]
    steps = np.diff(p)
    assert np.all(np.abs(steps - g) < 0.05 * g), steps
    faces = w.part(abi.PART_FIELD)[4 * 2 * W * H:4 * (2 * W * H + (W + 1) * H + W * (H + 1))].reshape(-1, 4)
    held = faces[:, 2] == 2.0
    assert np.max(np.abs(faces[held, 0])) < 0.02


def test_classify_bubble_and_gap():
    """isObstacleCell's bubble (fluid.wgsl:1123-1147): a cell of air with
    liquid beside it on two sides and two cells of liquid over it, pressing,
    is an obstacle to the liquid, so it holds its shape at the first step."""
    W, H = 16, 16
    cells = scenes.still_tank(W, H, depth=12)
    m = np.zeros((H, W), bool)
    m[10, 7] = True                  # a cell of air inside the water
    scenes.paint(cells, W, H, m, E.AIR)
    words = cells[:4 * W * H].reshape(H, W, 4)
    gas = cells[4 * W * H:].reshape(-1, 2)[:W * H].reshape(H, W, 2)
    gas[10, 7, 1] = np.uint32(np.float16(1.0).view(np.uint16)) << 16
    w = make_oracle(W, H, cells)
    sim = Simulation(w, W, H)
    sim.step()
    f = w.part(abi.PART_FIELD)[:8 * W * H].reshape(H, W, 8)
    # its first step the water above it is new (no falling bit): a bubble
    assert f[10, 7, 2] < 0.0
    assert f[9, 7, 2] > 0.0 and f[11, 7, 2] > 0.0


def test_pressure_solve_matches_fp64_cg():
    """The step's MG-PCG solution against the FP64 CG oracle on the same
    matrix, solved to 1e-12."""
    W, H = 40, 40
    w = make_oracle(W, H, scenes.dam_break(W, H, block=(20, 30)))
    sim = config_sim(w, W, H, tol=1e-7)
    for _ in range(3):
        sim.step()
    w.capture(True)
    sim.step()
    b = w.captured_b()
    x64, it, rel = w.cg64(b, tol=1e-12)
    assert rel < 1e-10
    f = w.part(abi.PART_FIELD)[:8 * W * H].reshape(-1, 8)
    fluid = f[:, 2] > 0
    x = f[:, 0]
    err = np.linalg.norm(x[fluid] - x64[fluid]) / np.linalg.norm(x64[fluid])
    assert err < 1e-5, err
    # x is stored in f32 (the field's layout): even the FP64 solution rounded
    # to f32 leaves a true residual of some 1e-6 to 1e-5 of b's; the solve's
    # stays within a few times that floor
    floor = w.relres(b, x64.astype(np.float32))
    assert w.relres(b, x) < 5 * floor, (w.relres(b, x), floor)
    s = w.solve_stats()
    assert s.rr <= (1e-7) ** 2 * s.bb   # the recursive residual met the tolerance


def test_fp64_cg_oracle_manufactured():
    """The FP64 CG oracle itself: in a closed box full of water (one region,
    pinned at its first cell), A x for a known x, solved back."""
    W, H = 12, 10
    cells = scenes.empty_cells(W, H)
    scenes.paint(cells, W, H, scenes.border(W, H), E.WALL)
    inner = ~scenes.border(W, H)
    scenes.paint(cells, W, H, inner, E.WATER)
    w = make_oracle(W, H, cells)
    Simulation(w, W, H, earth_g=0.0).step()
    f = w.part(abi.PART_FIELD)[:8 * W * H].reshape(H, W, 8)
    rho = f[:, :, 2]
    assert np.all(rho[inner] > 0)
    # A from the classified world: uniform water, share 1 everywhere inside
    rng = np.random.default_rng(5)
    xs = np.where(inner, rng.normal(size=(H, W)), 0.0)
    b = np.zeros((H, W))
    first = np.argwhere(inner)[0]
    for y, x in np.argwhere(inner):
        s = 0.0
        for dy, dx in ((0, -1), (0, 1), (-1, 0), (1, 0)):
            if inner[y + dy, x + dx]:
                s += xs[y, x] - xs[y + dy, x + dx]
        b[y, x] = s + (xs[y, x] if (y, x) == tuple(first) else 0.0)
    x64, it, rel = w.cg64(b.astype(np.float32).ravel(), tol=1e-13)
    assert np.allclose(x64.reshape(H, W)[inner], xs[inner], atol=1e-5)


@pytest.mark.parametrize("every", [1, 3])
@pytest.mark.parametrize("scene", ["dam", "precision"])
def test_residual_replacement_stops_on_the_true_residual(every, scene):
    """With residual replacement every k iterations (WigStepCtl.residual_every)
    the solve stops only on a recomputed residual, so its pressure's true
    ||b - A x|| / ||b|| (FP64, from the f32 x) meets the tolerance, or, where
    an f32 pressure cannot (the dam: even the FP64 CG solution rounded to
    f32 leaves 8e-6), comes within twice that floor.  The iterations run
    are a multiple of k."""
    W = H = 64
    tol = 1e-6
    if scene == "dam":
        w = make_oracle(W, H, scenes.dam_break(W, H, block=(40, 44)))
        sim = config_sim(w, W, H, tol=tol)
        warm = 3
    else:
        w = make_oracle(W, H, scenes.precision_world(W, 3)[0])
        sim = config_sim(w, W, H, tol=tol)
        sim.earth_g = 0.0
        warm = 0
    sim.solver = SolverSettings.relative(tol, residual_every=every)
    for _ in range(warm):
        sim.step()
    w.capture(True)
    sim.step()
    b = w.captured_b()
    x = w.part(abi.PART_FIELD)[:8 * W * H].reshape(-1, 8)[:, 0].copy()
    s = w.solve_stats()
    assert s.converged and s.cycles % every == 0, (s.cycles, every)
    rel = w.relres(b, x)
    x64, _, _ = w.cg64(b, tol=1e-13)
    floor = w.relres(b, x64.astype(np.float32))
    assert rel <= max(tol * 1.001, 2.2 * floor), (rel, floor)
    if scene == "precision":
        assert floor < 0.2 * tol and rel <= tol * 1.001
    assert abs(np.sqrt(s.rr / s.bb) - rel) <= 1e-3 * rel     # the stop's residual is the true one


def test_openmp_build_equals_serial():
    """The CPU baseline's build (OpenMP over the reference's independent
    invocations, atomics where they meet) steps every world to the serial
    oracle's buffers: cells, counters, field, blocks and fluidCg bit for bit,
    the particles as multisets (their slots' order is the atomics')."""
    import os
    from oracle.oracle import OracleWorld, lib, OMP_PATH
    from tests.helpers import compare_worlds
    omp = lib(OMP_PATH)
    W, H = 96, 80
    for cells in (scenes.dam_break(W, H, block=(48, 60)), scenes.discs_dam(W, H, discs=10, rmin=3, rmax=8)[0]):
        ws = OracleWorld(W, H, kinds=E.kinds())
        wp = OracleWorld(W, H, kinds=E.kinds(), L=omp)
        sims = []
        for w in (ws, wp):
            w.upload(abi.PART_CELLS, cells)
            sims.append(config_sim(w, W, H, tol=1e-6))
        for step in range(12):
            for s in sims:
                s.step()
            d = compare_worlds(ws, wp)
            assert not d, f"step {step}: " + "; ".join(d)
