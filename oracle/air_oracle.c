/*
 * The air's projection restated on the CPU: the MG-PCG that solves for the
 * air's dynamic pressure (shaders/air.wgsl:869-1536, dispatched by
 * Simulation.encodeAir, js/sim.js:2264-2289).  TEST INFRASTRUCTURE ONLY:
 * tests/, the smoke check and the benchmark's CPU baseline call it as the
 * checker; the product (libwig.so) never loads it.
 *
 * Buffers in the reference's layouts:
 *   mg      airProjection (js/sim.js:712-723): every level, two vec4s a
 *           cell, level l at levelBase(l) (air.wgsl:172-192):
 *           (potential, rhs, right weight, bottom weight),
 *           (edge conductance, fixed, piece bits, open share);
 *   cg      airCg (js/sim.js:858-864): per level-0 cell (x, r, p, A p), then
 *           the sums cgSumAt(k) (air.wgsl:951-961): r.z, alpha, beta, a
 *           spare, then one partial per 8 x 8 group of level 0;
 *   status  airFastest's words (air.wgsl:96-104): [1..3] the iterations'
 *           most |r| slots in turn, [4] the iterations run.
 *   mean    per level-0 cell, meanExpansion of its zone (air.wgsl:1006-1021),
 *           the caller's (its zones are outside the projection).
 *
 * The V-cycle is restated as whole-level passes: red-black sweeps over every
 * cell of a level, then the restriction of every coarse cell.  The reference
 * sweeps tiles with halos instead and states that they leave each cell as
 * such passes would, to the bit (air.wgsl:1199-1213); the HIP path sweeps
 * tiles of its own shape, so parity between the two checks that claim too.
 * Float arithmetic in the WGSL's order, no fused multiply-adds
 * (-ffp-contract=off), correctly rounded division.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define AIR_LEVELS 9        /* js/sim.js:28 */
#define AIR_COARSEST 256    /* js/shared.js:239 */
#define AIR_CYCLES 64       /* js/sim.js:36 */
#define ANCHOR 1e-6f        /* air.wgsl:146 */
#define COARSE_SWEEPS 20    /* air.wgsl:153 */
#define COARSE_OMEGA 1.7f   /* air.wgsl:154 */
#define COARSEN 1.8f        /* air.wgsl:889 */
#define CORRECT 1.5f        /* air.wgsl:901 */
#define AIR_TOLERANCE 1e-4f /* air.wgsl:945 */
#define AIR_SMOOTH 2        /* air.wgsl:1214 */

typedef struct {
    uint32_t L, w[AIR_LEVELS], h[AIR_LEVELS];
    uint64_t base[AIR_LEVELS]; /* levelBase, in vec4s */
} Air;

/* The levels: halve (rounding up) until a level has AIR_COARSEST cells or
 * fewer, AIR_LEVELS at most (js/sim.js:712-719); levelBase (air.wgsl:180-188). */
static void air_init(Air *a, uint32_t W, uint32_t H) {
    a->L = 1;
    a->w[0] = W;
    a->h[0] = H;
    while (a->L < AIR_LEVELS) {
        uint32_t lw = a->w[a->L - 1], lh = a->h[a->L - 1];
        if ((uint64_t)lw * lh <= AIR_COARSEST) break;
        a->w[a->L] = (lw + 1u) / 2u;
        a->h[a->L] = (lh + 1u) / 2u;
        a->L++;
    }
    uint64_t b = 0;
    for (uint32_t l = 0; l < a->L; l++) {
        a->base[l] = 2u * b;
        b += (uint64_t)a->w[l] * a->h[l];
    }
}

uint32_t wo_air_levels(uint32_t W, uint32_t H, uint32_t *lw, uint32_t *lh) {
    Air a;
    air_init(&a, W, H);
    for (uint32_t l = 0; l < a.L; l++) {
        if (lw) lw[l] = a.w[l];
        if (lh) lh[l] = a.h[l];
    }
    return a.L;
}

/* mg's and cg's sizes in floats (js/sim.js:720-723,858-864). */
uint64_t wo_air_mg_floats(uint32_t W, uint32_t H) {
    Air a;
    air_init(&a, W, H);
    uint64_t n = 0;
    for (uint32_t l = 0; l < a.L; l++) n += (uint64_t)a.w[l] * a.h[l];
    return 8u * n;
}
static uint32_t air_groups(uint32_t W, uint32_t H) { return ((W + 7u) / 8u) * ((H + 7u) / 8u); }
uint64_t wo_air_cg_floats(uint32_t W, uint32_t H) {
    return 4u * ((uint64_t)W * H + (4u + air_groups(W, H) + 3u) / 4u);
}

/* word k (0..7) of level l's cell (x, y): mgAt (air.wgsl:190-192) */
#define MG(l, x, y, k) mg[4u * (a->base[l] + 2u * ((uint64_t)(y) * a->w[l] + (uint64_t)(x))) + (k)]
#define HELD(l, x, y) (MG(l, x, y, 5) > 0.0f)
#define SUMS(k) cg[4u * ((uint64_t)a->w[0] * a->h[0]) + (k)]

static float f32bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t u32bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float coarsenAt(uint32_t l) { return l >= 1u ? COARSEN : 1.0f; }   /* :903-905 */

/* :912-916 */
static float anchorAt(const Air *a, uint32_t l, int32_t cx, int32_t cy) {
    int32_t lox = cx << l, loy = cy << l;
    int32_t hx = lox + (1 << l), hy = loy + (1 << l);
    if (hx > (int32_t)a->w[0]) hx = (int32_t)a->w[0];
    if (hy > (int32_t)a->h[0]) hy = (int32_t)a->h[0];
    return ANCHOR * (float)((hx - lox) * (hy - loy)) / coarsenAt(l);
}

/* converged (:963-965): the last iteration's most |r| under the tolerance */
static int converged(const uint32_t *status, uint32_t cycle) {
    return f32bits(status[1u + (cycle + 2u) % 3u]) < AIR_TOLERANCE;
}
static void maxBits(uint32_t *slot, float r) {
    uint32_t b = u32bits(fabsf(r));
    if (b > *slot) *slot = b;
}

/* applyA (:969-993) on cg's word `which` */
static float applyA(const Air *a, const float *mg, const float *cg, int32_t x, int32_t y, uint32_t which) {
    const uint32_t dx = a->w[0], dy = a->h[0];
    const uint64_t i = (uint64_t)y * dx + (uint64_t)x;
    float diag = MG(0, x, y, 2) + MG(0, x, y, 3) + MG(0, x, y, 4) + anchorAt(a, 0u, x, y);
    float sum = 0.0f;
    if ((uint32_t)x + 1u < dx) sum += MG(0, x, y, 2) * cg[4u * (i + 1u) + which];
    if ((uint32_t)y + 1u < dy) sum += MG(0, x, y, 3) * cg[4u * (i + dx) + which];
    if (x > 0) {
        float w = MG(0, x - 1, y, 2);
        sum += w * cg[4u * (i - 1u) + which];
        diag += w;
    }
    if (y > 0) {
        float w = MG(0, x, y - 1, 3);
        sum += w * cg[4u * (i - dx) + which];
        diag += w;
    }
    return diag * cg[4u * i + which] - sum;
}

/* gather's clearing of the solve's slots (air.wgsl:444-449) */
void wo_air_clear_status(uint32_t *status) {
    status[1] = 0u;
    status[3] = 0u;
    status[4] = 0u;
}

/* cgStart (:1027-1048) */
void wo_air_start(uint32_t W, uint32_t H, float *mg, float *cg, uint32_t *status, const float *mean) {
    Air A, *a = &A;
    air_init(a, W, H);
    const uint64_t n = (uint64_t)W * H;
    float *r = (float *)malloc(n * sizeof(float));
    /* every cell's r from the potentials as they were (a pass writes no
     * potential): then the writes */
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            const uint64_t i = (uint64_t)y * W + x;
            float rr = 0.0f;
            if (!HELD(0, x, y)) {
                if (MG(0, x, y, 7) > 0.0f) MG(0, x, y, 1) += mean ? mean[i] : 0.0f;
                rr = -MG(0, x, y, 1) - applyA(a, mg, cg, (int32_t)x, (int32_t)y, 0u);
            }
            r[i] = rr;
        }
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            const uint64_t i = (uint64_t)y * W + x;
            cg[4u * i + 1u] = r[i];
            cg[4u * i + 2u] = 0.0f;
            cg[4u * i + 3u] = 0.0f;
            MG(0, x, y, 0) = 0.0f;
            MG(0, x, y, 1) = -r[i];
            maxBits(&status[3], r[i]);
        }
    free(r);
}

/* A level's cells as a tile of the V-cycle holds them (loadTile,
 * :1255-1291): the correction, rhs, right and bottom weights, the diagonal
 * (-1 where not swept: fixed).  On the way down the correction starts at
 * none; on the way back (`up`) at the one down left (preAt, :1236-1241)
 * plus CORRECT times the coarser level's (but on a fixed cell). */
typedef struct { float *x, *rhs, *right, *down, *diag; } Level;

static void loadLevel(const Air *a, const float *mg, const float *cg, uint32_t l, int up, Level *t) {
    const uint32_t dx = a->w[l], dy = a->h[l];
    for (uint32_t y = 0; y < dy; y++)
        for (uint32_t x = 0; x < dx; x++) {
            const uint64_t k = (uint64_t)y * dx + x;
            float cx = 0.0f, diag = -1.0f;
            if (up) cx = l == 0 ? cg[4u * k + 3u] : MG(l, x, y, 6);
            if (!(MG(l, x, y, 5) > 0.0f)) {
                if (up) cx += CORRECT * MG(l + 1u, x / 2u, y / 2u, 0);
                diag = MG(l, x, y, 2) + MG(l, x, y, 3) + MG(l, x, y, 4) + anchorAt(a, l, (int32_t)x, (int32_t)y);
                if (x > 0) diag += MG(l, x - 1u, y, 2);
                if (y > 0) diag += MG(l, x, y - 1u, 3);
            }
            t->x[k] = cx;
            t->rhs[k] = MG(l, x, y, 1);
            t->right[k] = MG(l, x, y, 2);
            t->down[k] = MG(l, x, y, 3);
            t->diag[k] = diag;
        }
}

/* tileSum (:1293-1309) */
static float levelSum(const Level *t, uint32_t dx, uint32_t dy, uint32_t x, uint32_t y) {
    const uint64_t k = (uint64_t)y * dx + x;
    float sum = 0.0f;
    if (x + 1u < dx) sum += t->right[k] * t->x[k + 1u];
    if (y + 1u < dy) sum += t->down[k] * t->x[k + dx];
    if (x > 0) sum += t->right[k - 1u] * t->x[k - 1u];
    if (y > 0) sum += t->down[k - dx] * t->x[k - dx];
    return sum;
}

/* sweepTile (:1313-1326) over the whole level: the cells of one colour,
 * each from its neighbours (all of the other colour) */
static void sweepLevel(Level *t, uint32_t dx, uint32_t dy, int color) {
    for (uint32_t y = 0; y < dy; y++)
        for (uint32_t x = 0; x < dx; x++) {
            const uint64_t k = (uint64_t)y * dx + x;
            if ((int)((x + y) & 1u) != color || t->diag[k] < 0.0f) continue;
            t->x[k] = (levelSum(t, dx, dy, x, y) - t->rhs[k]) / t->diag[k];
        }
}

/* restrictCell (:1349-1417): level l + 1's cell (cx, cy) from its four */
static void restrictCell(const Air *a, float *mg, const Level *t, uint32_t l, int32_t cx, int32_t cy) {
    const int32_t fdx = (int32_t)a->w[l], fdy = (int32_t)a->h[l];
    float rhs = 0.0f, edge = 0.0f;
    uint32_t nfree = 0, held = 0;
    for (int q = 0; q < 4; q++) {
        const int32_t fx = 2 * cx + (q & 1), fy = 2 * cy + (q >> 1);
        if (fx >= fdx || fy >= fdy) continue;
        if (HELD(l, fx, fy)) {
            held++;
            continue;
        }
        nfree++;
        edge += MG(l, fx, fy, 4);
        if (fx + 1 < fdx && HELD(l, fx + 1, fy)) edge += MG(l, fx, fy, 2);
        if (fy + 1 < fdy && HELD(l, fx, fy + 1)) edge += MG(l, fx, fy, 3);
        if (fx > 0 && HELD(l, fx - 1, fy)) edge += MG(l, fx - 1, fy, 2);
        if (fy > 0 && HELD(l, fx, fy - 1)) edge += MG(l, fx, fy - 1, 3);
        const uint64_t k = (uint64_t)fy * (uint32_t)fdx + (uint32_t)fx;
        rhs += t->rhs[k] - (levelSum(t, (uint32_t)fdx, (uint32_t)fdy, (uint32_t)fx, (uint32_t)fy) - t->diag[k] * t->x[k]);
    }
    const int32_t right = 2 * cx + 1 < fdx - 1 ? 2 * cx + 1 : fdx - 1;
    const int32_t bottom = 2 * cy + 1 < fdy - 1 ? 2 * cy + 1 : fdy - 1;
    float wR = 0.0f, wD = 0.0f;
    for (int r = 0; r < 2; r++) {
        const int32_t y = 2 * cy + r;
        if (y < fdy && !HELD(l, right, y) && (right + 1 >= fdx || !HELD(l, right + 1, y))) wR += MG(l, right, y, 2);
        const int32_t x = 2 * cx + r;
        if (x < fdx && !HELD(l, x, bottom) && (bottom + 1 >= fdy || !HELD(l, x, bottom + 1))) wD += MG(l, x, bottom, 3);
    }
    const int fixed = nfree == 0u && held > 0u;
    const uint32_t lc = l + 1u;
    const float scale = lc == 1u ? 1.0f / COARSEN : 1.0f;
    MG(lc, cx, cy, 0) = 0.0f;
    MG(lc, cx, cy, 1) = fixed ? 0.0f : rhs;
    MG(lc, cx, cy, 2) = wR * scale;
    MG(lc, cx, cy, 3) = wD * scale;
    MG(lc, cx, cy, 4) = edge * scale;
    MG(lc, cx, cy, 5) = (float)fixed;
    MG(lc, cx, cy, 6) = 0.0f;
    MG(lc, cx, cy, 7) = 0.0f;
}

static void levelAlloc(Level *t, uint64_t n) {
    float *p = (float *)malloc(5u * n * sizeof(float));
    t->x = p;
    t->rhs = p + n;
    t->right = p + 2u * n;
    t->down = p + 3u * n;
    t->diag = p + 4u * n;
}

/* smoothDown (:1428-1447): `down` (restricts) or `smoothTile` */
static void smoothDown(const Air *a, float *mg, float *cg, uint32_t l, int restricts) {
    const uint32_t dx = a->w[l], dy = a->h[l];
    Level t;
    levelAlloc(&t, (uint64_t)dx * dy);
    loadLevel(a, mg, cg, l, 0, &t);
    for (int i = 0; i < AIR_SMOOTH; i++) {
        sweepLevel(&t, dx, dy, 0);
        sweepLevel(&t, dx, dy, 1);
    }
    /* storeTile (:1328-1339) */
    for (uint32_t y = 0; y < dy; y++)
        for (uint32_t x = 0; x < dx; x++) {
            const uint64_t k = (uint64_t)y * dx + x;
            if (!(t.diag[k] >= 0.0f)) continue;
            if (!restricts) MG(l, x, y, 0) = t.x[k];
            else if (l == 0) cg[4u * k + 3u] = t.x[k];
            else MG(l, x, y, 6) = t.x[k];
        }
    if (restricts)
        for (uint32_t cy = 0; cy < a->h[l + 1u]; cy++)
            for (uint32_t cx = 0; cx < a->w[l + 1u]; cx++) restrictCell(a, mg, &t, l, (int32_t)cx, (int32_t)cy);
    free(t.x);
}

/* up (:1461-1474) */
static void levelUp(const Air *a, float *mg, const float *cg, uint32_t l) {
    const uint32_t dx = a->w[l], dy = a->h[l];
    Level t;
    levelAlloc(&t, (uint64_t)dx * dy);
    loadLevel(a, mg, cg, l, 1, &t);
    for (int i = 0; i < AIR_SMOOTH; i++) {
        sweepLevel(&t, dx, dy, 1);
        sweepLevel(&t, dx, dy, 0);
    }
    for (uint32_t y = 0; y < dy; y++)
        for (uint32_t x = 0; x < dx; x++) {
            const uint64_t k = (uint64_t)y * dx + x;
            if (t.diag[k] >= 0.0f) MG(l, x, y, 0) = t.x[k];
        }
    free(t.x);
}

/* solveCoarsest (:1482-1535): symmetric SOR in one workgroup's memory */
static void solveCoarsest(const Air *a, float *mg, uint32_t l) {
    const uint32_t dx = a->w[l], dy = a->h[l];
    uint32_t n = dx * dy;
    if (n > AIR_COARSEST) n = AIR_COARSEST;
    float A[AIR_COARSEST][4], B[AIR_COARSEST][2];
    for (uint32_t j = 0; j < n; j++) {
        for (int k = 0; k < 4; k++) A[j][k] = mg[4u * (a->base[l] + 2u * j) + k];
        for (int k = 0; k < 2; k++) B[j][k] = mg[4u * (a->base[l] + 2u * j + 1u) + k];
    }
    for (int sweep = 0; sweep < 2 * COARSE_SWEEPS; sweep++) {
        const int color = sweep < COARSE_SWEEPS ? (sweep & 1) : ((2 * COARSE_SWEEPS - 1 - sweep) & 1);
        for (uint32_t j = 0; j < n; j++) {
            const int32_t cx = (int32_t)(j % dx), cy = (int32_t)(j / dx);
            if (((cx + cy) & 1) != color || B[j][1] > 0.0f) continue;
            float sum = 0.0f;
            float diag = A[j][2] + A[j][3] + B[j][0] + anchorAt(a, l, cx, cy);
            if ((uint32_t)cx + 1u < dx) sum += A[j][2] * A[j + 1u][0];
            if ((uint32_t)cy + 1u < dy) sum += A[j][3] * A[j + dx][0];
            if (cx > 0) {
                sum += A[j - 1u][2] * A[j - 1u][0];
                diag += A[j - 1u][2];
            }
            if (cy > 0) {
                sum += A[j - dx][3] * A[j - dx][0];
                diag += A[j - dx][3];
            }
            /* mix(e1, e2, t) = e1 (1 - t) + e2 t */
            const float x0 = A[j][0], xn = (sum - A[j][1]) / diag;
            A[j][0] = x0 * (1.0f - COARSE_OMEGA) + xn * COARSE_OMEGA;
        }
    }
    for (uint32_t j = 0; j < n; j++) mg[4u * (a->base[l] + 2u * j)] = A[j][0];
}

/* One V-cycle on level 0's rhs (js/sim.js:2270-2282): each level but the
 * coarsest smoothed and restricted, the coarsest solved, each level back up
 * corrected and smoothed; level 0's answer in its potential words. */
static void vcycle(const Air *a, float *mg, float *cg) {
    const uint32_t deepest = a->L - 1u;
    for (uint32_t l = 0; l < deepest; l++) smoothDown(a, mg, cg, l, 1);
    if (deepest == 0) smoothDown(a, mg, cg, 0, 0);
    solveCoarsest(a, mg, deepest);
    for (int32_t l = (int32_t)deepest - 1; l >= 0; l--) levelUp(a, mg, cg, (uint32_t)l);
}
void wo_air_vcycle(uint32_t W, uint32_t H, float *mg, float *cg) {
    Air A;
    air_init(&A, W, H);
    vcycle(&A, mg, cg);
}

/* sumGroup (:1054-1069): the 8 x 8 tree of a group's values */
static float groupTree(float v[64]) {
    for (int s = 32; s > 0; s >>= 1)
        for (int li = 0; li < s; li++) v[li] += v[li + s];
    return v[0];
}
/* sumAll (:1081-1097): 256 lanes, each its strided share, then the tree */
static float sumAll(const Air *a, const float *cg) {
    const uint32_t n = air_groups(a->w[0], a->h[0]);
    float wide[256];
    for (uint32_t li = 0; li < 256u; li++) {
        float s = 0.0f;
        for (uint32_t k = li; k < n; k += 256u) s += SUMS(4u + k);
        wide[li] = s;
    }
    for (int h = 128; h > 0; h >>= 1)
        for (int li = 0; li < h; li++) wide[li] += wide[li + h];
    return wide[0];
}
/* a partial per group of level 0 (cgRho's r z, cgApply's p A p) */
static void groupSums(const Air *a, float *cg, const float *v) {
    const uint32_t dx = a->w[0], dy = a->h[0], gx = (dx + 7u) / 8u, gy = (dy + 7u) / 8u;
    for (uint32_t by = 0; by < gy; by++)
        for (uint32_t bx = 0; bx < gx; bx++) {
            float p[64];
            for (uint32_t li = 0; li < 64u; li++) {
                const uint32_t x = bx * 8u + li % 8u, y = by * 8u + li / 8u;
                p[li] = (x < dx && y < dy) ? v[(uint64_t)y * dx + x] : 0.0f;
            }
            SUMS(4u + by * gx + bx) = groupTree(p);
        }
}

/* One iteration `cycle` (js/sim.js:2270-2288): the V-cycle, cgRho (:1099-1112),
 * cgBeta (:1114-1126), cgDirection (:1128-1139), cgApply (:1141-1157),
 * cgAlpha (:1159-1170), cgUpdate (:1172-1194); each but cgUpdate's slot
 * clearing skipped once converged. */
void wo_air_iteration(uint32_t W, uint32_t H, float *mg, float *cg, uint32_t *status, uint32_t cycle) {
    Air A, *a = &A;
    air_init(a, W, H);
    const uint64_t n = (uint64_t)W * H;
    const int conv = converged(status, cycle);
    if (!conv) {
        vcycle(a, mg, cg);
        float *v = (float *)malloc(n * sizeof(float));
        for (uint32_t y = 0; y < H; y++)
            for (uint32_t x = 0; x < W; x++) v[(uint64_t)y * W + x] = cg[4u * ((uint64_t)y * W + x) + 1u] * MG(0, x, y, 0);
        groupSums(a, cg, v);
        {   /* cgBeta */
            const float rho = sumAll(a, cg), before = SUMS(0);
            SUMS(2) = (cycle == 0u || before == 0.0f) ? 0.0f : rho / before;
            SUMS(0) = rho;
        }
        for (uint32_t y = 0; y < H; y++)   /* cgDirection */
            for (uint32_t x = 0; x < W; x++) {
                const uint64_t i = (uint64_t)y * W + x;
                cg[4u * i + 2u] = HELD(0, x, y) ? 0.0f : MG(0, x, y, 0) + SUMS(2) * cg[4u * i + 2u];
            }
        for (uint32_t y = 0; y < H; y++)   /* cgApply */
            for (uint32_t x = 0; x < W; x++) {
                const uint64_t i = (uint64_t)y * W + x;
                v[i] = HELD(0, x, y) ? 0.0f : applyA(a, mg, cg, (int32_t)x, (int32_t)y, 2u);
            }
        for (uint64_t i = 0; i < n; i++) {
            cg[4u * i + 3u] = v[i];
            v[i] = cg[4u * i + 2u] * v[i];
        }
        groupSums(a, cg, v);
        {   /* cgAlpha */
            const float pq = sumAll(a, cg);
            SUMS(1) = pq <= 0.0f ? 0.0f : SUMS(0) / pq;
        }
        free(v);
    }
    /* cgUpdate */
    status[1u + (cycle + 1u) % 3u] = 0u;
    if (conv) return;
    status[4] += 1u;
    const float alpha = SUMS(1);
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            const uint64_t i = (uint64_t)y * W + x;
            float *v = &cg[4u * i];
            const float r = v[1] - alpha * v[3];
            v[0] = v[0] + alpha * v[2];
            v[1] = r;
            MG(0, x, y, 0) = 0.0f;
            MG(0, x, y, 1) = -r;
            maxBits(&status[1u + cycle % 3u], r);
        }
}

/* The projection's solve as encodeAir dispatches it: the slots cleared,
 * cgStart, then `cycles` iterations (those after convergence skip). */
void wo_air_project(uint32_t W, uint32_t H, float *mg, float *cg, uint32_t *status, const float *mean, uint32_t cycles) {
    wo_air_clear_status(status);
    wo_air_start(W, H, mg, cg, status, mean);
    for (uint32_t k = 0; k < cycles; k++) wo_air_iteration(W, H, mg, cg, status, k);
}
