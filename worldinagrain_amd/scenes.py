"""Initial worlds as the reference's cells buffer (js/shared.js:60-79):
16 bytes a cell (x word, temperature bits, velocity as 2 x f16, w), then the
gas plane, 8 bytes a cell (vapour and mist, 2 x f16; oxygen used up and amount
of gas, 2 x f16).  Liquid is painted undrawn, so the step's syncSeed seeds it
with LIQUID_PER_CELL particles a cell (fluid.wgsl:863-890), as after a brush
stroke.  Worlds are y-down; the scene sizes are given y-up and flipped."""
import numpy as np

from . import elements as E


def _f32bits(x):
    return np.frombuffer(np.float32(x).tobytes(), dtype=np.uint32)[0]


def empty_cells(width, height, temp=20.0):
    """A world of air at `temp`, each cell holding a cell of air at 1 atm."""
    C = width * height
    words = np.zeros(4 * C + 4 * ((C + 1) // 2), dtype=np.uint32)
    cells = words[:4 * C].reshape(C, 4)
    cells[:, 0] = E.AIR
    cells[:, 1] = _f32bits(temp)
    gas = words[4 * C:].reshape(-1, 2)[:C]           # a (vapour|mist, lost|amount) pair a cell
    gas[:, 1] = np.uint32(np.float16(1.0).view(np.uint16)) << 16
    return words


def paint(words, width, height, mask, elem, temp=20.0, vel=None):
    """Cells where mask (H x W, y-down) is true become `elem` (undrawn)."""
    C = width * height
    cells = words[:4 * C].reshape(height, width, 4)
    gas = words[4 * C:].reshape(-1, 2)[:C].reshape(height, width, 2)
    cells[mask, 0] = elem
    cells[mask, 1] = _f32bits(temp)
    cells[mask, 3] = 0
    if vel is None:
        cells[mask, 2] = 0
    else:
        vx = np.asarray(vel[0], dtype=np.float16).view(np.uint16).astype(np.uint32)
        vy = np.asarray(vel[1], dtype=np.float16).view(np.uint16).astype(np.uint32)
        packed = vx | (vy << 16)
        cells[mask, 2] = packed[mask] if packed.shape == mask.shape else packed
    gas[mask] = 0
    return words


def border(width, height, thickness=1):
    m = np.zeros((height, width), dtype=bool)
    m[:thickness, :] = m[-thickness:, :] = True
    m[:, :thickness] = m[:, -thickness:] = True
    return m


def dam_break(width=256, height=256, block=(128, 128), wall=1):
    """A solid border `wall` cells thick and water filling the cells
    [wall, block) x [wall, block), y-up from the bottom left."""
    w = empty_cells(width, height)
    paint(w, width, height, border(width, height, wall), E.WALL)
    m = np.zeros((height, width), dtype=bool)
    for yu in range(wall, block[1]):
        m[height - 1 - yu, wall:block[0]] = True
    paint(w, width, height, m, E.WATER)
    return w


def still_tank(width=32, height=32, depth=16, wall=1):
    """A closed box of water `depth` rows deep at rest."""
    w = empty_cells(width, height)
    paint(w, width, height, border(width, height, wall), E.WALL)
    m = np.zeros((height, width), dtype=bool)
    m[height - wall - depth:height - wall, wall:width - wall] = True
    paint(w, width, height, m, E.WATER)
    return w


class SplitMix64:
    def __init__(self, seed):
        self.s = np.uint64(seed)

    def next(self):
        with np.errstate(over="ignore"):
            self.s = np.uint64(self.s + np.uint64(0x9E3779B97F4A7C15))
            z = self.s
            z = np.uint64((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9))
            z = np.uint64((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB))
            return int(z ^ (z >> np.uint64(31)))

    def uniform(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


def discs_dam(width=1024, height=1024, discs=200, rmin=4, rmax=24, solid_share=0.10, fluid_rows=0.40,
              vel_sigma=0.5, dt=1.0 / 120.0, seed=2):
    """Random solid discs, radius U[rmin, rmax], centres uniform (SplitMix64
    `seed`), up to `discs` of them or until `solid_share` of the cells are
    solid; water in the non-solid cells of the lowest `fluid_rows` of the
    rows, each cell's velocity N(0, vel_sigma^2) cells per unit time per
    component, times dt per step (all particles seeded in a cell share it)."""
    rng = SplitMix64(seed)
    solid = np.zeros((height, width), dtype=bool)
    yy, xx = np.mgrid[0:height, 0:width]
    for _ in range(discs):
        if solid.mean() >= solid_share:
            break
        cx = rng.uniform() * width
        cy = rng.uniform() * height
        r = rmin + rng.uniform() * (rmax - rmin)
        solid |= (xx + 0.5 - cx) ** 2 + (yy + 0.5 - cy) ** 2 <= r * r
    w = empty_cells(width, height)
    paint(w, width, height, solid, E.WALL)
    fluid = np.zeros((height, width), dtype=bool)
    rows = int(round(fluid_rows * height))
    fluid[height - rows:, :] = True
    fluid &= ~solid
    g = np.random.default_rng(seed)
    vel = (g.normal(0.0, vel_sigma, size=(height, width)) * dt, g.normal(0.0, vel_sigma, size=(height, width)) * dt)
    paint(w, width, height, fluid, E.WATER, vel=vel)
    return w, int(fluid.sum())


def batch_world(side, seed):
    """A world of the batching workload: a side x side grid with solid
    discs over a share U[0, 0.15] of its cells (radius U[2, side / 16 + 2]),
    and a dam-break block of water at rest over a share U[0.2, 0.5] of them:
    a rectangle standing on the floor against the left wall, its width a
    random U[sqrt(share), 1] of the side (SplitMix64 `seed`).  Returns
    (cells, water cells)."""
    rng = SplitMix64(seed)
    share = 0.2 + 0.3 * rng.uniform()
    solid_share = 0.15 * rng.uniform()
    solid = np.zeros((side, side), dtype=bool)
    yy, xx = np.mgrid[0:side, 0:side]
    rmax = side / 16.0 + 2.0
    for _ in range(400):
        if solid.mean() >= solid_share:
            break
        cx, cy = rng.uniform() * side, rng.uniform() * side
        r = 2.0 + rng.uniform() * (rmax - 2.0)
        solid |= (xx + 0.5 - cx) ** 2 + (yy + 0.5 - cy) ** 2 <= r * r
    a = np.sqrt(share) + (1.0 - np.sqrt(share)) * rng.uniform()
    bw = max(1, min(side, int(round(a * side))))
    bh = max(1, min(side, int(round(share * side * side / bw))))
    w = empty_cells(side, side)
    paint(w, side, side, solid, E.WALL)
    fluid = np.zeros((side, side), dtype=bool)
    fluid[side - bh:, :bw] = True
    fluid &= ~solid
    paint(w, side, side, fluid, E.WATER)
    return w, int(fluid.sum())


def batch_sides(n, seed=7, sides=(64, 128, 256, 512), probs=(0.4, 0.3, 0.2, 0.1)):
    """n world sides drawn with the given probabilities (SplitMix64 `seed`)."""
    rng = SplitMix64(seed)
    cum = np.cumsum(probs)
    out = []
    for _ in range(n):
        u = rng.uniform()
        out.append(sides[int(min(np.searchsorted(cum, u, side="right"), len(sides) - 1))])
    return out


def precision_world(side, seed, fluid_share=0.70, air_share=0.20, solid_share=0.10, vel_sigma=1.0, dt=1.0 / 120.0):
    """A world of the precision study: solid discs over `solid_share` of the
    cells (radius U[2, side / 64 + 4]), then air in random axis-aligned
    rectangles (sides U[side / 32, side / 8]) over `air_share`, the rest
    water (about `fluid_share`) with each cell's velocity i.i.d.
    N(0, vel_sigma^2) per component, cells per unit time times dt
    (SplitMix64 and numpy `seed`).  Returns (cells, water cells)."""
    rng = SplitMix64(seed)
    solid = np.zeros((side, side), dtype=bool)
    yy, xx = np.mgrid[0:side, 0:side]
    rmax = side / 64.0 + 4.0
    while solid.mean() < solid_share:
        cx, cy = rng.uniform() * side, rng.uniform() * side
        r = 2.0 + rng.uniform() * (rmax - 2.0)
        x0, x1 = max(0, int(cx - r) - 1), min(side, int(cx + r) + 2)
        y0, y1 = max(0, int(cy - r) - 1), min(side, int(cy + r) + 2)
        solid[y0:y1, x0:x1] |= (xx[y0:y1, x0:x1] + 0.5 - cx) ** 2 + (yy[y0:y1, x0:x1] + 0.5 - cy) ** 2 <= r * r
    air = np.zeros((side, side), dtype=bool)
    lo, hi = max(1, side // 32), max(2, side // 8)
    while (air & ~solid).mean() < air_share:
        w = lo + int(rng.uniform() * (hi - lo))
        h = lo + int(rng.uniform() * (hi - lo))
        x0, y0 = int(rng.uniform() * (side - w)), int(rng.uniform() * (side - h))
        air[y0:y0 + h, x0:x0 + w] = True
    cells = empty_cells(side, side)
    paint(cells, side, side, solid, E.WALL)
    fluid = ~solid & ~air
    g = np.random.default_rng(seed)
    vel = (g.normal(0.0, vel_sigma, size=(side, side)) * dt, g.normal(0.0, vel_sigma, size=(side, side)) * dt)
    paint(cells, side, side, fluid, E.WATER, vel=vel)
    return cells, int(fluid.sum())


def air_rooms(width, height, seed=1, walls=12, liquid_share=0.15, holes=2, rhs_sigma=0.05, warm_sigma=0.01, flow_seed=None,
              fractional=False):
    """Level 0 of the air's projection as gather fills it (shaders/air.wgsl:
    352-414, 825-866), for a world of walls and liquid in air.

    Returns (mg0, x0): mg0 a (height, width, 8) float32 array of level 0's two
    vec4s a cell, x0 the warm-start potential (cg's x).  Walls are random
    axis-aligned rectangles (every one a hollow room: sealed pockets of air
    inside, held to none only by ANCHOR); the lower `liquid_share` of rows
    holds liquid (no gas: faces shut, its open share -1).  A face is open
    (weight 1) where the cells across it both carry gas; the last column's
    right faces, the last row's bottom faces and the first column's left
    edges are open where their cell carries gas (the world's open edges).
    `holes` black holes are held at a potential below none (fixed).  The
    right-hand side is the net outflow of face velocities N(0, rhs_sigma) on
    the open faces (drawn from `flow_seed` if given: the same rooms, another
    step's flow); the warm start N(0, warm_sigma) in cells that carry gas.
    With `fractional`, the open faces' weights are drawn from (0.25, 1]."""
    rng = np.random.default_rng(seed)
    W, H = width, height
    carries = np.ones((H, W), bool)
    for _ in range(walls):
        w, h = int(rng.integers(6, max(7, W // 3))), int(rng.integers(6, max(7, H // 3)))
        x0, y0 = int(rng.integers(0, max(1, W - w))), int(rng.integers(0, max(1, H - h)))
        carries[y0:y0 + h, x0:x0 + w] = False
        carries[y0 + 1:y0 + h - 1, x0 + 1:x0 + w - 1] = True   # a room inside
        if rng.random() < 0.5 and w > 4 and y0 + h // 2 < H:  # a door in some
            carries[y0 + h // 2, x0] = True
    liquid = np.zeros((H, W), bool)
    liquid[int(H * (1.0 - liquid_share)):, :] = True
    liquid &= carries
    carries &= ~liquid
    fixed = np.zeros((H, W), bool)
    ys, xs = np.nonzero(carries)
    if len(ys):
        for j in rng.choice(len(ys), size=min(holes, len(ys)), replace=False):
            fixed[ys[j], xs[j]] = True
    right = np.zeros((H, W), bool)
    right[:, :-1] = carries[:, :-1] & carries[:, 1:]
    right[:, -1] = carries[:, -1]
    down = np.zeros((H, W), bool)
    down[:-1, :] = carries[:-1, :] & carries[1:, :]
    down[-1, :] = carries[-1, :]
    edge = np.zeros((H, W), bool)
    edge[:, 0] = carries[:, 0]
    share = np.where(carries, 1.0, np.where(liquid, -1.0, 0.0)).astype(np.float32)
    mg0 = np.zeros((H, W, 8), np.float32)
    # divergenceAt's net outflow (:825-838) of random face velocities: zero
    # summed over a sealed room, as the flow through its faces cancels
    flow = rng if flow_seed is None else np.random.default_rng(flow_seed)
    u = flow.normal(0.0, rhs_sigma, (H, W)).astype(np.float32) * right
    v = flow.normal(0.0, rhs_sigma, (H, W)).astype(np.float32) * down
    out = u + v
    out[:, 1:] -= u[:, :-1]
    out[1:, :] -= v[:-1, :]
    mg0[..., 1] = out
    if fractional:
        # open faces of any weight in (0.25, 1] (gather's are 1: the solve's
        # arithmetic is the same for any)
        mg0[..., 2] = right * rng.uniform(0.25, 1.0, (H, W))
        mg0[..., 3] = down * rng.uniform(0.25, 1.0, (H, W))
        mg0[..., 4] = edge * rng.uniform(0.25, 1.0, (H, W))
        mg0[..., 1] = out * 0.5
    else:
        mg0[..., 2] = right
        mg0[..., 3] = down
        mg0[..., 4] = edge
    mg0[..., 5] = fixed
    mg0[..., 6] = np.where(carries, np.float32(0.0), np.uint32(0xffffffff).view(np.float32))   # piece bits
    mg0[..., 7] = share
    x0 = np.where(carries, rng.normal(0.0, warm_sigma, (H, W)), 0.0).astype(np.float32)
    x0[fixed] = -0.5                                           # -sinkSpeed (gather, :406-408)
    return mg0, x0


def air_over(width, height, words, seed=1, wind=0.05, temp=20.0, lapse=0.002):
    """An air[0] buffer (H x W x 4 f32: pressure, the right and bottom
    faces' velocities, temperature) over the world `words`, as the air's own
    step would hand the liquid step one: a pressure rising down the world
    by `lapse` a row about 1 atm plus noise, winds N(0, wind) on every face,
    temperatures about `temp`; liquid cells flagged AIR_BLOCKED and solid
    cells AIR_SOLID (airsample.wgsl:130-132)."""
    from . import abi
    rng = np.random.default_rng(seed)
    C = width * height
    kinds = E.kinds()
    elem = (words[:4 * C].reshape(height, width, 4)[..., 0] & 0xFF).astype(np.int64)
    kind = np.asarray(kinds, dtype=np.int64)[elem]
    air = np.empty((height, width, 4), dtype=np.float32)
    rows = np.arange(height, dtype=np.float32)[:, None]
    air[..., 0] = 12.0 + lapse * rows + rng.normal(0.0, lapse, (height, width))
    air[..., 1:3] = rng.normal(0.0, wind, (height, width, 2))
    air[..., 3] = temp + rng.normal(0.0, 1.0, (height, width))
    air[kind == E.KIND_LIQUID, 3] = abi.AIR_BLOCKED
    air[(kind == E.KIND_SOLID) | (kind == E.KIND_POWDER), 3] = abi.AIR_SOLID
    return air


class Body:
    """A body as the bodies' pass leaves it in its table (js/shared.js:316-326's
    BODY_WORD words): its stamped cells (a y-down boolean mask), its pose's
    centre `pos` (cells), velocity, spin, mass and moment of inertia (cells of
    water; None: `density` times its cells, and a rectangle's about its
    centre), whether it is held, its pose's subcell offset (SUBCELL), and a
    mobility (six words of a symmetric 3 x 3) taken when `mobility_at` is
    the step before the one run (MOBILITY_AT)."""

    def __init__(self, slot, mask, density=1.0, pos=None, vel=(0.0, 0.0), spin=0.0, mass=None, inertia=None, held=False,
                 wet=False, subcell=(0.0, 0.0), mobility=None, mobility_at=0xFFFFFFFF, cargo=0.0, cargo_inertia=0.0,
                 alive=True):
        self.slot, self.mask = slot, np.asarray(mask, dtype=bool)
        ys, xs = np.nonzero(self.mask)
        n = len(xs)
        self.cells = n
        if pos is None:
            pos = (float(xs.mean()) + 0.5, float(ys.mean()) + 0.5)
        self.pos = pos
        self.vel, self.spin = vel, spin
        self.mass = density * n if mass is None else mass
        if inertia is None:
            inertia = float(self.mass / max(n, 1) * (((xs + 0.5 - pos[0]) ** 2 + (ys + 0.5 - pos[1]) ** 2).sum() + n / 6.0))
        self.inertia = inertia
        self.held, self.wet, self.alive = held, wet, alive
        self.subcell = subcell
        self.mobility, self.mobility_at = mobility, mobility_at
        self.cargo, self.cargo_inertia = cargo, cargo_inertia


def _f16pair(a, b):
    return int(np.float16(a).view(np.uint16)) | (int(np.float16(b).view(np.uint16)) << 16)


def body_parts(width, height, words, bodies, elem=E.STONE, cups=(), seams=(), frame=1):
    """The parts of a world with free and held bodies: each body's cells
    painted `elem` marked loose (FL_LOOSE) with its slot in frac.x (common.wgsl:
    156-162), its table slot in the motion part, and the swept words of the
    step before `frame` for the cells of `cups` and `seams` ((body slot, mask)
    pairs; common.wgsl:355-370).  Returns (motion, swept) as u32 arrays."""
    from . import abi
    C = width * height
    cells = words[:4 * C].reshape(height, width, 4)
    motion = np.zeros(abi.MOTION_WORDS, dtype=np.uint32)
    mf = motion.view(np.float32)
    for b in bodies:
        paint(words, width, height, b.mask, elem)
        cells[b.mask, 0] |= np.uint32(abi.FL_LOOSE << 28)
        cells[b.mask, 3] = _f16pair(float(b.slot), 0.0)
        base = abi.body_word(b.slot, 0)
        mf[base + abi.BW_POS], mf[base + abi.BW_POS + 1] = b.pos
        mf[base + abi.BW_VEL], mf[base + abi.BW_VEL + 1] = b.vel
        mf[base + abi.BW_SPIN] = b.spin
        mf[base + abi.BW_MASS] = b.mass
        mf[base + abi.BW_INERTIA] = b.inertia
        motion[base + abi.BW_CELLS] = b.cells
        motion[base + abi.BW_FLAGS] = ((abi.BF_ALIVE if b.alive else 0) | (abi.BF_HELD if b.held else 0)
                                       | (abi.BF_WET if b.wet else 0))
        motion[base + abi.BW_SUBCELL] = _f16pair(*b.subcell)
        mf[base + abi.BW_CARGO] = b.cargo
        mf[base + abi.BW_CARGO_INERTIA] = b.cargo_inertia
        if b.mobility is not None:
            mf[base + abi.BW_MOBILITY:base + abi.BW_MOBILITY + 6] = b.mobility
        motion[base + abi.BW_MOBILITY_AT] = b.mobility_at & 0xFFFFFFFF
    swept = np.zeros(2 * C, dtype=np.uint32)
    back = (frame - 1) & 0xFFFFFFFF
    for seam, marks in ((False, cups), (True, seams)):
        for slot, mask in marks:
            idx = np.flatnonzero(np.asarray(mask, dtype=bool).ravel())
            swept[2 * idx + (back & 1)] = abi.swept_key(back, slot, seam)
    return motion, swept


def bodies_tank(width=64, height=48, depth=28, wall=1, frame=1):
    """A closed tank of water `depth` rows deep with bodies in it: a light
    floating plank, a heavy box sinking and turning (its mobility given for
    the first step, its pose a quarter cell off its stamp), a held block on
    the floor, wet, and a cup (a U of a free body) with powder in it and a
    seam in its wall.  Returns (cells, motion, swept, kinds, props)."""
    w = still_tank(width, height, depth, wall)
    surface = height - wall - depth
    def rect(x0, y0, x1, y1):
        m = np.zeros((height, width), dtype=bool)
        m[y0:y1, x0:x1] = True
        return m
    plank = rect(6, surface - 1, 18, surface + 2)
    box = rect(24, surface + 8, 29, surface + 13)
    block = rect(36, height - wall - 4, 44, height - wall)
    cup = rect(48, surface + 4, 58, surface + 12) & ~rect(50, surface + 4, 56, surface + 10)
    inside = rect(50, surface + 4, 56, surface + 10)
    seam = rect(49, surface + 6, 50, surface + 7)
    cup_wall = cup & ~seam
    bodies = [
        Body(3, plank, density=0.5, vel=(0.01, 0.0)),
        Body(7, box, density=2.0, vel=(0.0, 0.05), spin=0.01, subcell=(0.25, -0.125),
             mobility=(0.02, 0.001, 0.0005, 0.03, 0.0002, 0.004), mobility_at=frame - 1),
        Body(12, block, density=2.5, held=True, wet=True),
        Body(20, cup_wall, density=0.8, cargo=4.0, cargo_inertia=10.0),
    ]
    motion, swept = body_parts(width, height, w, bodies, cups=[(20, inside)], seams=[(20, seam)], frame=frame)
    # powder settled in the cup's floor: its cargo (classify, :1616-1619)
    sand = rect(50, surface + 9, 56, surface + 10)
    paint(w, width, height, sand, E.SAND)
    kinds = E.kinds()
    props = np.zeros((256, 6), dtype=np.float32)
    props[E.STONE, 5] = 1.0
    return w, motion, swept, kinds, props
