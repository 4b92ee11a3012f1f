// The liquid step's kernels for gfx950, in the order js/sim.js encodeFluid
// dispatches the reference's passes (js/sim.js:2466-2612).  Each kernel
// cites the pass of shaders/fluid.wgsl it computes.
//
// Launch shapes: a workgroup of 64 threads is one wavefront and one 8 x 8
// block of cells (li = y * 8 + x), and the passes over the blocks worked on
// run a fixed grid that strides over the device-side active list (its count
// is never read back to the host).  The passes over particles stride over
// `count` the same way, 256 threads a workgroup.  Sums that must come out
// the same every run are integer (fixed point), or FP64 reduced in one
// fixed tree: a wavefront's 64 cells by halving, then the blocks' partials
// in block-index order by one workgroup.
#include "wig_device.h"

#define WAVE 64
#define WAVES 4   // wavefronts a workgroup of 256

__device__ __forceinline__ uint32_t ld(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The 16-bit V-cycle's rhs scale: the power of two that takes a residual
// whose largest |r| has these float bits into [0.5, 1) (the PCG's iterates
// do not depend on the preconditioner's scale; 16 bits do not reach the
// residuals of a converging solve unscaled).  1 for float storage.
__device__ __forceinline__ float vcycleScale(const Dev &d, uint32_t most) {
    uint32_t e = (most >> 23) & 0xffu;
    if (d.store == 0u || most == 0u || e == 0xffu) return 1.0f;
    return __uint_as_float((253u - e) << 23);
}
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
// An iteration pass's words, read together (one round trip to memory, not
// one a test): the iteration's number; whether the world's host would have
// encoded it (within its fluid_cycles, and before the convergence
// checkpoint its solve stops at: a batch runs as many iterations as its
// slowest world, the others' skip); whether the solve had converged before
// it; and the active blocks' count.
struct Iter { uint32_t cycle, active; bool encoded, converged; };
__device__ __forceinline__ Iter iterWords(const Dev &d) {
    const uint32_t cycle = ld(&d.wstate[WS_CYCLE]), stop = ld(&d.wstate[WS_STOP]);
    const uint32_t conv = __hip_atomic_load(&d.head->converged, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t active = d.blocks[COUNT_SLOT(d, COUNT_ACTIVE)];
    Iter it;
    it.cycle = cycle;
    it.active = active;
    it.encoded = d.full && cycle < d.fluid_cycles && cycle < stop;
    it.converged = cycle > 0u && conv != 0u;
    return it;
}

// ------------------------------------------------------------ cell tests
__device__ __forceinline__ float densityAt(const Dev &d, i2 c) {           // fluid.wgsl:322-327
    if (!inWorld(d, c)) return 0.0f;
    return (float)CNT(d, cellIndex(d, c), G_DENSITY) / (FIXF * (float)LIQUID_PER_CELL);
}
__device__ __forceinline__ float gridZ(const Dev &d, i2 c) { return inWorld(d, c) ? FLD(d, 2u * cellIndex(d, c), 2) : -1.0f; }
__device__ __forceinline__ float gridX(const Dev &d, i2 c) { return inWorld(d, c) ? FLD(d, 2u * cellIndex(d, c), 0) : 0.0f; }
__device__ __forceinline__ float driftAt(const Dev &d, i2 c) { return inWorld(d, c) ? FLD(d, 2u * cellIndex(d, c) + 1u, 0) : 0.0f; }
__device__ __forceinline__ float faceBeta(float a, float b) {              // :369-374
    if (a < 0.0f || b < 0.0f || (a <= 0.0f && b <= 0.0f)) return 0.0f;
    return (a > 0.0f && b > 0.0f) ? 2.0f / (a + b) : 1.0f / fmaxf(a, b);
}
__device__ __forceinline__ float surfaceShare(float da, float db) {       // :384-386
    return wclamp((da - 0.5f) / fmaxf(da - db, 1e-3f), THETA_MIN, 1.0f);
}
__device__ __forceinline__ float fineBeta(float a, float b, float da, float db) {   // :396-405
    float beta = faceBeta(a, b);
    if (a > 0.0f && b == 0.0f) return beta / (da > 0.0f ? surfaceShare(da, db) : 0.5f);
    if (b > 0.0f && a == 0.0f) return beta / (db > 0.0f ? surfaceShare(db, da) : 0.5f);
    return beta;
}
__device__ __forceinline__ bool isLiquidAt(const Dev &d, i2 c) {
    return inWorld(d, c) && kindOf(d, wType(d.cells[4ull * cellIndex(d, c)])) == KIND_LIQUID;
}
__device__ __forceinline__ uint32_t stepBit(const Dev &d, uint32_t back) { return 1u << ((d.frame + back) & 1u); }
__device__ __forceinline__ uint32_t fallingBit(const Dev &d, uint32_t back) { return 4u << ((d.frame + back) & 1u); }
__device__ __forceinline__ bool fallingAt(const Dev &d, i2 q) {
    return inWorld(d, q) && (wu32(FLD(d, 2u * cellIndex(d, q) + 1u, 3)) & fallingBit(d, 1u)) != 0u;
}
// ------------------------------------------------------------ the air
// air[0] and airSource as the liquid reads and adds to them (AIR_CELL 1: an
// air cell is a world cell)
__device__ __forceinline__ bool inAir(const Dev &d, i2 i) { return i.x >= 0 && i.y >= 0 && i.x < d.W && i.y < d.H; }
__device__ __forceinline__ float4 airAt(const Dev &d, i2 i) { return d.air[(uint64_t)i.y * (uint32_t)d.W + (uint32_t)i.x]; }
__device__ __forceinline__ v2 fluidPressureAt(const Dev &d, i2 i) {            // airsample.wgsl:83-89
    if (!inAir(d, i)) return V2(0.0f, 0.0f);
    const float4 a = airAt(d, i);
    return V2(a.x, a.w <= AIR_SOLID ? 0.0f : 1.0f);
}
__device__ float pressureSlope(const Dev &d, i2 c, i2 s) {                     // airsample.wgsl:97-111
    const float own = airAt(d, c).x;
    const v2 lo = fluidPressureAt(d, I2(c.x - s.x, c.y - s.y)), hi = fluidPressureAt(d, I2(c.x + s.x, c.y + s.y));
    if (lo.y > 0.0f && hi.y > 0.0f) return 0.5f * (hi.x - lo.x);
    if (hi.y > 0.0f) return hi.x - own;
    if (lo.y > 0.0f) return own - lo.x;
    return 0.0f;
}
__device__ v2 airGrad(const Dev &d, v2 pos) {                                  // sampleAir's grad, :171-180
    const v2 q = V2(pos.x / 1.0f, pos.y / 1.0f);
    int32_t cx = wi32(floorf(q.x)), cy = wi32(floorf(q.y));
    cx = cx < 0 ? 0 : (cx > d.W - 1 ? d.W - 1 : cx);
    cy = cy < 0 ? 0 : (cy > d.H - 1 ? d.H - 1 : cy);
    return V2(pressureSlope(d, I2(cx, cy), I2(1, 0)) / 1.0f, pressureSlope(d, I2(cx, cy), I2(0, 1)) / 1.0f);
}
__device__ __forceinline__ v2 pressurePush(const Dev &d, v2 grad, float rho) {  // airsample.wgsl:11-14
    const float s = d.dt / d.dx, m = fmaxf(rho, 1e-3f);
    return V2(-grad.x * (101325.0f / P_ATM) * s * s / m, -grad.y * (101325.0f / P_ATM) * s * s / m);
}
__device__ __forceinline__ bool gasCellAt(const Dev &d, i2 i) { return inAir(d, i) && airAt(d, i).w > AIR_BLOCKED; }
__device__ __forceinline__ float faceU(const Dev &d, i2 i) { return inAir(d, i) ? airAt(d, i).y : 0.0f; }
__device__ __forceinline__ float faceV(const Dev &d, i2 i) { return inAir(d, i) ? airAt(d, i).z : 0.0f; }
__device__ void addSource(const Dev &d, uint32_t i, uint32_t k, float x) {     // airsource.wgsl:27-35
    if (x == 0.0f) return;
    const uint64_t at = SOURCE_WORDS * (uint64_t)i + 2u * k;
    const u2 p = exactParts(x, EXACT_FINE);
    const uint32_t old = atomicAdd(&d.source[at], p.lo);
    atomicAdd(&d.source[at + 1u], p.hi + (old + p.lo < old ? 1u : 0u));
}
__device__ void pushFace(const Dev &d, i2 i, v2 dv) {                          // airsource.wgsl:37-41,59-63
    if (!inAir(d, i)) return;
    const uint32_t j = (uint32_t)i.y * (uint32_t)d.W + (uint32_t)i.x;
    addSource(d, j, SOURCE_RIGHT, dv.x);
    addSource(d, j, SOURCE_DOWN, dv.y);
}
__device__ __forceinline__ v2 pairDrag(float u, float wv, float ml, float ma, float c) {   // fluid.wgsl:1869-1878
    if (ml <= 0.0f || ma <= 0.0f || c <= 0.0f) return V2(0.0f, 0.0f);
    const float mu = ml / ma;
    const float k = c * fabsf(u - wv);
    const float change = (u - wv) * (1.0f - 1.0f / (1.0f + k * (1.0f + 1.0f / mu)));
    return V2(-change / (1.0f + mu), change * mu / (1.0f + mu));
}
__device__ __forceinline__ float gasAmountAt(const Dev &d, uint32_t i) {
    uint64_t at = gasWord(d, i);
    v2 m = unpack2x16(d.cells[at]);
    v2 r = unpack2x16(d.cells[at + 1]);
    return r.y + m.x * (AMBIENT_T + KELVIN) / (BOIL_K * VAPOUR_MAX);
}
__device__ float airMassAt(const Dev &d, i2 c) { return AIR_RHO / 1000.0f * gasAmountAt(d, cellIndex(d, c)); }   // :1862-1865
// the wind's drag along the surface at face (a | b) (surfaceDrag, :1914-1932)
__device__ float surfaceDrag(const Dev &d, i2 a, i2 b, bool isU, float4 ga, float4 gb, float u) {
    const i2 across = isU ? I2(0, 1) : I2(1, 0);
    float du = 0.0f;
    if (ga.z > 0.0f && gb.z > 0.0f) {
        for (int s = -1; s <= 1; s += 2) {
            const i2 oa = I2(a.x + across.x * s, a.y + across.y * s), ob = I2(b.x + across.x * s, b.y + across.y * s);
            if (gasCellAt(d, oa) && gasCellAt(d, ob)) {
                const float wv = isU ? faceU(d, oa) : faceV(d, oa);
                const v2 drag = pairDrag(u + du, wv, 0.5f * (ga.z + gb.z), 0.5f * (airMassAt(d, oa) + airMassAt(d, ob)), d.surface_drag);
                du += drag.x;
                pushFace(d, oa, isU ? V2(drag.y, 0.0f) : V2(0.0f, drag.y));
            }
        }
    }
    return du;
}
// people's bodies (common.wgsl:238-254): cell i's occupant word, 0 for
// nobody or with occupants off
#define OCCUPANT_SCALE 128.0f
__device__ __forceinline__ uint32_t occupantAt(const Dev &d, uint32_t i) { return d.occ_on ? d.occ[i] : 0u; }
__device__ __forceinline__ bool isOccupant(uint32_t w) { return (w & 0xffu) != 0u; }
__device__ __forceinline__ v2 occupantVel(uint32_t w) {
    // (the 12-bit fields sign-extended by shifting them to the top first)
    const int32_t x = (int32_t)(w << 12u) >> 20, y = (int32_t)w >> 20;
    return V2((float)x / OCCUPANT_SCALE, (float)y / OCCUPANT_SCALE);
}
// The pull and down (pull.wgsl:13-44) from the gravity pass's field, a
// vec4 a light cell of GRAVITY_BLOCK x GRAVITY_BLOCK cells: the masses' pull
// (xy), their potential (z), down (w).  Off, Earth's pull alone and down
// (0, 1), as the field's texel for that (no masses' pull, down step 2).
#define GRAVITY_BLOCK 2
__device__ __forceinline__ float4 gravityTexel(const Dev &d, i2 c) {
    if (!d.grav_on) return make_float4(0.0f, 0.0f, 0.0f, 2.0f);
    const int32_t gw = (d.W + 1) / 2, gh = (d.H + 1) / 2;
    const i2 k = clampTo(c, gw, gh);
    return d.grav[(uint32_t)k.y * (uint32_t)gw + (uint32_t)k.x];
}
// fieldAt's pull (xy): bilinear between the light cells' middles
__device__ __forceinline__ v2 fieldPull(const Dev &d, v2 pos) {
    const float qx = pos.x / (float)GRAVITY_BLOCK - 0.5f, qy = pos.y / (float)GRAVITY_BLOCK - 0.5f;
    const float fx0 = floorf(qx), fy0 = floorf(qy);
    const i2 i = I2(wi32(fx0), wi32(fy0));
    const float fx = qx - fx0, fy = qy - fy0;
    const float4 a = gravityTexel(d, i), b = gravityTexel(d, I2(i.x + 1, i.y));
    const float4 c = gravityTexel(d, I2(i.x, i.y + 1)), e = gravityTexel(d, I2(i.x + 1, i.y + 1));
    const v2 top = V2(wmix(a.x, b.x, fx), wmix(a.y, b.y, fx)), bottom = V2(wmix(c.x, e.x, fx), wmix(c.y, e.y, fx));
    return V2(wmix(top.x, bottom.x, fy), wmix(top.y, bottom.y, fy));
}
__device__ __forceinline__ v2 pullAt(const Dev &d, v2 pos) {
    if (!d.grav_on) return V2(0.0f, d.earthG);
    const v2 f = fieldPull(d, pos);
    return V2(0.0f + f.x, d.earthG + f.y);
}
__device__ __forceinline__ i2 downAt(const Dev &d, i2 p) {      // (p in the world)
    if (!d.grav_on) return I2(0, 1);
    return stepAt(wu32(gravityTexel(d, I2(p.x / GRAVITY_BLOCK, p.y / GRAVITY_BLOCK)).w) & 7u);
}
// eighths of a turn from the Earth's down to the down at c (:3822-3825)
__device__ __forceinline__ int32_t downTurns(const Dev &d, i2 c) {
    return (int32_t)stepIndex(downAt(d, clampTo(c, d.W, d.H))) - 2;
}
__device__ bool isObstacleCell(const Dev &d, i2 c, uint32_t vx) {       // :1123-1147
    if (inWorld(d, c) && isOccupant(occupantAt(d, cellIndex(d, c)))) return true;
    uint32_t t = wType(vx), k = kindOf(d, t);
    if (k == KIND_GAS) {
        int wet = 0;
        for (int j = 0; j < 4; j++) wet += isLiquidAt(d, beside(c, j)) ? 1 : 0;
        if (wet < 2) return false;
        if (t == E_AIR) {
            const i2 dn = downAt(d, c);
            i2 up = I2(c.x - dn.x, c.y - dn.y), up2 = I2(c.x - 2 * dn.x, c.y - 2 * dn.y);
            if (!isLiquidAt(d, up) || !isLiquidAt(d, up2) || fallingAt(d, up)) return false;
            return gasAmountAt(d, cellIndex(d, c)) >= BUBBLE_GAS;
        }
        return true;
    }
    return k == KIND_SOLID || k == KIND_POWDER;
}
__device__ int32_t countAt(const Dev &d, i2 c) {                           // :1150-1159
    if (!inWorld(d, c)) return -1;
    uint32_t i = cellIndex(d, c);
    if (isObstacleCell(d, c, d.cells[4ull * i])) return -1;
    return (int32_t)CNT(d, i, G_COUNT);
}
__device__ __forceinline__ float4 gridAt4(const Dev &d, i2 c) {
    if (!inWorld(d, c)) return make_float4(0.0f, 0.0f, -1.0f, 0.0f);
    return *(const float4 *)&FLD(d, 2u * cellIndex(d, c), 0);
}
// ------------------------------------------------------------ bodies
// Free and held bodies in the liquid (fluid.wgsl:1181-1475): the bodies'
// table in `motion` (js/shared.js:294-333's BODY_WORD words), the bodies
// buffer's swept words (common.wgsl:355-370), and the step's push sums.
// With the bodies off (d.body_on 0) none is read and no cell is a body's.
struct v3 { float x, y, z; };
__device__ __forceinline__ v3 V3(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ uint32_t bodyWord(const Dev &d, uint32_t b, uint32_t k) {    // :1211-1213
    return d.motion[MAX_BODIES + b * BODY_TABLE_WORDS + k];
}
__device__ __forceinline__ float bodyF(const Dev &d, uint32_t b, uint32_t k) { return __uint_as_float(bodyWord(d, b, k)); }
__device__ __forceinline__ v2 bodyV(const Dev &d, uint32_t b, uint32_t k) { return V2(bodyF(d, b, k), bodyF(d, b, k + 1u)); }
__device__ __forceinline__ uint32_t liquidSlot(uint32_t b, uint32_t w) {                  // :1226-1228
    return MAX_BODIES * (1u + BODY_TABLE_WORDS) + BODY_CONTACTS * BODY_CONTACT_WORDS + b * BODY_LIQUID_WORDS + w;
}
// alive and weighed (:1244-1246)
__device__ __forceinline__ bool bodyWeighed(const Dev &d, uint32_t b) {
    const uint32_t flags = bodyWord(d, b, BW_FLAGS);
    return !((flags & BF_ALIVE) == 0u || (flags & BF_NEW) != 0u || bodyF(d, b, BW_MASS) <= 0.0f);
}
__device__ uint32_t rigidBodyOf(const Dev &d, uint4 v) {                                 // :1233-1248
    if (!d.body_on) return WNONE;
    const uint32_t t = wType(v.x);
    if (!wHas(v.x, FL_LOOSE) || kindOf(d, t) != KIND_SOLID || d.eprops[t].breaks == 0.0f) return WNONE;
    const float id = unpack2x16(v.w).x;
    if (id < 0.0f || id >= (float)MAX_BODIES) return WNONE;
    const uint32_t b = wu32(id);
    return bodyWeighed(d, b) ? b : WNONE;
}
__device__ __forceinline__ uint32_t claimStep(uint32_t frame) { return frame % CLAIM_PERIOD + 1u; }   // common.wgsl:384-386
// the body whose cup (seam 0) or seam (1) cell i was stamped as `back` steps
// ago, alive and weighed when `weighed` (cupOf :1252-1266, seamOf
// :1272-1286; cupBefore :3520-3527 asks only for a slot)
__device__ uint32_t sweptBody(const Dev &d, uint32_t i, uint32_t back, uint32_t seam, bool weighed) {
    if (!d.body_on) return WNONE;
    const uint32_t frame = d.frame - back;
    const uint32_t key = d.swept[2ull * i + (frame & 1u)];
    if ((key >> 12) != claimStep(frame) || (key & 1u) != seam) return WNONE;
    const uint32_t b = 2047u - ((key >> 1) & 0x7ffu);
    if (b >= MAX_BODIES) return WNONE;
    if (weighed && !bodyWeighed(d, b)) return WNONE;
    return b;
}
__device__ __forceinline__ uint32_t cupOf(const Dev &d, uint32_t i) { return sweptBody(d, i, 1u, 0u, true); }
__device__ __forceinline__ uint32_t seamOf(const Dev &d, uint32_t i) { return sweptBody(d, i, 1u, 1u, true); }
__device__ __forceinline__ uint32_t cupBefore(const Dev &d, uint32_t i) { return sweptBody(d, i, 2u, 0u, false); }
__device__ __forceinline__ float rigidMass(const Dev &d, uint32_t b) { return bodyF(d, b, BW_MASS) + bodyF(d, b, BW_CARGO); }
__device__ __forceinline__ float rigidInertia(const Dev &d, uint32_t b) { return bodyF(d, b, BW_INERTIA) + bodyF(d, b, BW_CARGO_INERTIA); }
__device__ __forceinline__ float bodyDensity(const Dev &d, uint32_t b) {                 // :1300-1302
    return rigidMass(d, b) / fmaxf((float)bodyWord(d, b, BW_CELLS), 1.0f);
}
__device__ __forceinline__ uint32_t pushSlot(uint32_t b, uint32_t which, uint32_t k) { return b * BODY_PUSH_WORDS + 6u * which + 2u * k; }
__device__ void addPush(const Dev &d, uint32_t b, uint32_t which, v3 v) {                 // :1324-1331
    const float c[3] = {v.x, v.y, v.z};
#pragma unroll
    for (uint32_t k = 0; k < 3u; k++) {
        const u2 e = exactParts(c[k], EXACT_COARSE);
        const uint32_t at = pushSlot(b, which, k);
        const uint32_t old = atomicAdd(&d.pushes[at], e.lo);
        atomicAdd(&d.pushes[at + 1u], e.hi + (old + e.lo < old ? 1u : 0u));
    }
}
__device__ v3 pushOf(const Dev &d, uint32_t b, uint32_t which) {                           // :1333-1340
    float c[3];
#pragma unroll
    for (uint32_t k = 0; k < 3u; k++) {
        const uint32_t at = pushSlot(b, which, k);
        u2 e;
        e.lo = d.pushes[at];
        e.hi = d.pushes[at + 1u];
        c[k] = exactValue(e, EXACT_COARSE);
    }
    return V3(c[0], c[1], c[2]);
}
__device__ v3 giveWay(const Dev &d, uint32_t b, v3 g) {                                     // :1349-1356
    if (bodyWord(d, b, BW_MOBILITY_AT) != d.frame - 1u)
        return V3(g.x / rigidMass(d, b), g.y / rigidMass(d, b), g.z / rigidInertia(d, b));
    float m[6];
#pragma unroll
    for (uint32_t k = 0; k < 6u; k++) m[k] = bodyF(d, b, BW_MOBILITY + k);
    return V3(m[0] * g.x + m[1] * g.y + m[2] * g.z, m[1] * g.x + m[3] * g.y + m[4] * g.z, m[2] * g.x + m[4] * g.y + m[5] * g.z);
}
__device__ __forceinline__ v3 pushChange(const Dev &d, uint32_t b, uint32_t which) { return giveWay(d, b, pushOf(d, b, which)); }
__device__ __forceinline__ v3 pushChangeFree(const Dev &d, uint32_t b, uint32_t which) {   // :1367-1370
    const v3 g = pushOf(d, b, which);
    return V3(g.x / rigidMass(d, b), g.y / rigidMass(d, b), g.z / rigidInertia(d, b));
}
// the face from cell c to body b's cell q: (n, r x n) (:1374-1378)
__device__ v3 bodyFaceJ(const Dev &d, i2 c, i2 q, uint32_t b) {
    const v2 n = V2((float)(q.x - c.x), (float)(q.y - c.y));
    const v2 pos = bodyV(d, b, BW_POS);
    const v2 r = V2((float)c.x + 0.5f + 0.5f * n.x - pos.x, (float)c.y + 0.5f + 0.5f * n.y - pos.y);
    return V3(n.x, n.y, r.x * n.y - r.y * n.x);
}
// the body whose cell c is in the solve, as classify found it (:1422-1429),
// and the held body's (:1450-1457)
__device__ __forceinline__ uint32_t rigidAt(const Dev &d, i2 c) {
    if (!d.body_on || !inWorld(d, c)) return WNONE;
    const float z = FLD(d, 2u * cellIndex(d, c) + 1u, 2);
    const uint32_t b = wu32(fmaxf(-z, 1.0f)) - 1u;
    return z < 0.0f && b < MAX_BODIES ? b : WNONE;
}
__device__ __forceinline__ uint32_t heldAt(const Dev &d, i2 c) {
    if (!d.body_on || !inWorld(d, c)) return WNONE;
    const float z = FLD(d, 2u * cellIndex(d, c) + 1u, 2);
    const uint32_t b = wu32(fmaxf(-z, 1.0f)) - 1u;
    return z < 0.0f && b >= MAX_BODIES ? b - MAX_BODIES : WNONE;
}
__device__ __forceinline__ bool isHeld(const Dev &d, uint32_t b) { return (bodyWord(d, b, BW_FLAGS) & BF_HELD) != 0u; }
__device__ __forceinline__ v2 rigidVel(const Dev &d, uint32_t b, v2 x) {                   // :1460-1464
    const v2 pos = bodyV(d, b, BW_POS), vel = bodyV(d, b, BW_VEL);
    const v2 r = V2(x.x - pos.x, x.y - pos.y);
    const float spin = bodyF(d, b, BW_SPIN);
    return V2(vel.x + -spin * r.y, vel.y + spin * r.x);
}
__device__ __forceinline__ bool shutAt(const Dev &d, i2 c) { return gridZ(d, c) < 0.0f || rigidAt(d, c) != WNONE; }   // :1468-1470
__device__ __forceinline__ v2 faceMiddle(i2 at, bool isU) { return isU ? V2((float)at.x, (float)at.y + 0.5f) : V2((float)at.x + 0.5f, (float)at.y); }
__device__ __forceinline__ float bodyReach(const Dev &d, uint32_t b, i2 n) {                // :413-415
    const v2 s = unpack2x16(bodyWord(d, b, BW_SUBCELL));
    return wclamp(0.5f + (s.x * (float)n.x + s.y * (float)n.y), 0.0f, 1.0f);
}
__device__ void addPressurePush(const Dev &d, i2 c, float p, uint32_t which) {            // :1381-1389
    for (int k = 0; k < 4; k++) {
        const i2 q = beside(c, k);
        const uint32_t b = rigidAt(d, q);
        if (b != WNONE) {
            const v3 j = bodyFaceJ(d, c, q, b);
            addPush(d, b, which, V3(j.x * p, j.y * p, j.z * p));
        }
    }
}
__device__ float couplingAt(const Dev &d, i2 c, uint32_t which) {                          // :1393-1403
    float sum = 0.0f;
    for (int k = 0; k < 4; k++) {
        const i2 q = beside(c, k);
        const uint32_t b = rigidAt(d, q);
        if (b != WNONE) {
            const v3 j = bodyFaceJ(d, c, q, b), ch = pushChange(d, b, which);
            sum += j.x * ch.x + j.y * ch.y + j.z * ch.z;
        }
    }
    return sum;
}
// body b's velocity at x going into the solve (:1410-1419)
__device__ v2 bodyMoving(const Dev &d, uint32_t b, v2 x, bool drag) {
    const v2 pos = bodyV(d, b, BW_POS);
    const v2 pull = pullAt(d, pos);
    const float mass = rigidMass(d, b);
    const v3 s = pushChange(d, b, PUSH_SURFACE), g = giveWay(d, b, V3(mass * pull.x, mass * pull.y, 0.0f));
    v3 ch = V3(s.x + g.x, s.y + g.y, s.z + g.z);
    if (drag) {
        const v3 dr = pushChange(d, b, PUSH_DRAG);
        ch = V3(ch.x + dr.x, ch.y + dr.y, ch.z + dr.z);
    }
    const v2 r = V2(x.x - pos.x, x.y - pos.y);
    const v2 rv = rigidVel(d, b, x), px = pullAt(d, x);
    return V2(rv.x + px.x - pull.x + ch.x + -ch.z * r.y, rv.y + px.y - pull.y + ch.y + ch.z * r.x);
}
// the pull's potential at pos, Earth's and the masses' (:2954-2956)
__device__ float potentialAt(const Dev &d, v2 pos) {
    float z = 0.0f;
    if (d.grav_on) {
        const float qx = pos.x / (float)GRAVITY_BLOCK - 0.5f, qy = pos.y / (float)GRAVITY_BLOCK - 0.5f;
        const float fx0 = floorf(qx), fy0 = floorf(qy);
        const i2 i = I2(wi32(fx0), wi32(fy0));
        const float fx = qx - fx0, fy = qy - fy0;
        const float4 a = gravityTexel(d, i), b = gravityTexel(d, I2(i.x + 1, i.y));
        const float4 c = gravityTexel(d, I2(i.x, i.y + 1)), e = gravityTexel(d, I2(i.x + 1, i.y + 1));
        z = wmix(wmix(a.z, b.z, fx), wmix(c.z, e.z, fx), fy);
    }
    return d.earthG * ((float)d.H - pos.y) + z;
}
// the liquid's pressure seeping in at x on body h's face from c toward n
// (:2976-3001)
__device__ float seepAt(const Dev &d, i2 c, i2 n, uint32_t h, bool held, v2 x) {
    i2 ways[2] = {I2(-1, 0), I2(1, 0)};
    int count = 2;
    if (n.x != 0) {
        ways[0] = I2(-n.x, -n.y);
        count = 1;
    }
    int nearest = SEEP_REACH + 1;
    float p = 0.0f;
    for (int k2 = 0; k2 < count; k2++) {
        for (int k = 1; k < nearest; k++) {
            const i2 q = I2(c.x + ways[k2].x * k, c.y + ways[k2].y * k);
            if ((held ? heldAt(d, q) : rigidAt(d, q)) == h) continue;
            const float4 g = gridAt4(d, q);
            if (g.z > 0.0f) {
                nearest = k;
                p = g.x + g.z * (potentialAt(d, V2((float)q.x + 0.5f, (float)q.y + 0.5f)) - potentialAt(d, x));
            }
            break;
        }
    }
    return p;
}
// what the liquid's pressure on held body h's cell c does to it (:3010-3035)
__device__ void heldPush(const Dev &d, i2 c, uint32_t h, int32_t out[3]) {
    const v2 pos = bodyV(d, h, BW_POS);
    v2 f = V2(0.0f, 0.0f);
    float torque = 0.0f;
    for (int k = 0; k < 4; k++) {
        const i2 q = beside(c, k);
        const i2 n = I2(q.x - c.x, q.y - c.y);
        if (heldAt(d, q) == h) continue;
        const float4 g = gridAt4(d, q);
        const v2 x = V2((float)c.x + 0.5f + 0.5f * (float)n.x, (float)c.y + 0.5f + 0.5f * (float)n.y);
        float p = g.x;
        if (g.z > 0.0f) p += g.z * (potentialAt(d, V2((float)q.x + 0.5f, (float)q.y + 0.5f)) - potentialAt(d, x));
        else if (g.z < 0.0f) p = seepAt(d, c, n, h, true, x);
        const v2 push = V2(-p * (float)n.x, -p * (float)n.y);
        f = V2(f.x + push.x, f.y + push.y);
        const v2 r = V2(x.x - pos.x, x.y - pos.y);
        torque += r.x * push.y - r.y * push.x;
    }
    const float m = rigidMass(d, h), in = rigidInertia(d, h);
    out[0] = wi32(wround(f.x / m * RIGID_FIX));
    out[1] = wi32(wround(f.y / m * RIGID_FIX));
    out[2] = wi32(wround(torque / in * RIGID_FIX));
}
// a cell of a held body that was wet in the last solve (:583-595)
__device__ __forceinline__ bool wetRegionAt(const Dev &d, uint4 v) {
    const uint32_t b = rigidBodyOf(d, v);
    return b != WNONE && (bodyWord(d, b, BW_FLAGS) & (BF_WET | BF_HELD)) == (BF_WET | BF_HELD);
}
// e^x from basic float operations (round to nearest, no fused
// multiply-add): x = k ln2 + r, k = rint(x / ln2), ln2 split in two (Cody
// and Waite); e^r by its Taylor series to r^7 on |r| <= ln2 / 2; times 2^k
__device__ __forceinline__ float wexpf(float x) {
    if (!(x < 88.5f)) return x != x ? x : __int_as_float(0x7f800000);
    if (x < -103.0f) return 0.0f;
    const float k = rintf(x * 1.44269504f);
    const float r = (x - k * 0.693145752f) - k * 1.42860677e-6f;
    const float p = 1.0f + r * (1.0f + r * (0.5f + r * (0.166666672f + r * (0.0416666679f + r * (0.00833333377f
                    + r * (0.00138888892f + r * 0.000198412701f))))));
    return ldexpf(p, (int)k);
}
// viscosityAt (common.wgsl:1486-1493): a liquid's viscosity at temp
#define VISCOSITY_MOST 1e4f
__device__ __forceinline__ float viscosityAt(const WigElementProps &e, float temp) {
    if (e.viscosity_e <= 0.0f) return e.viscosity;
    return fminf(e.viscosity * wexpf((e.viscosity_t - temp) / e.viscosity_e), VISCOSITY_MOST);
}
// surface tension's constants (fluid.wgsl:1715-1732; js/shared.js:285-286)
// and waterTension (:1736-1739): water's σ in the solve's units, ρ g l²
#define CAPILLARY_LENGTH 0.0027f
#define SURFACE_MIN_LENGTH 0.3f
#define MAX_CURVATURE 1.0f
#define CURVATURE_GAIN 2.0f
#define CONTACT 2.0f
#define SURFACE_DAMP 1.0f
__device__ __forceinline__ float waterTension(const Dev &d) {
    const float l = CAPILLARY_LENGTH / d.dx;
    return l >= SURFACE_MIN_LENGTH ? WATER_WEIGHT * l * l : 0.0f;
}
__device__ v2 liquidIn(const Dev &d, uint32_t i) {                         // :1165-1179
    uint32_t vx = d.cells[4ull * i];
    float temp = fbits(d.cells[4ull * i + 1]);
    uint32_t k = kindOf(d, wType(vx));
    // (at the temperature it acts at: water's effTempIn, every liquid having
    // water's phase points here)
    const WigElementProps e = d.eprops[wType(vx)];
    float viscosity = k == KIND_LIQUID ? viscosityAt(e, effTemp(temp)) : 0.0f;
    uint32_t n = CNT(d, i, G_COUNT);
    if (n > 0u) return V2((float)CNT(d, i, G_MASS) / (FIXF * (float)n), viscosity);
    if (k != KIND_LIQUID) return V2(1.0f, 0.0f);
    return V2(liquidDensity(temp), viscosity);
}
__device__ __forceinline__ bool isObstaclePos(const Dev &d, v2 pos) {     // :424-431
    i2 c = I2(wi32(floorf(pos.x)), wi32(floorf(pos.y)));
    if (gridZ(d, c) < 0.0f) return true;
    // (outside the world gridZ is -1: the words below are never read there)
    if (isOccupant(occupantAt(d, cellIndex(d, c)))) return true;
    uint32_t k = kindOf(d, wType(d.cells[4ull * cellIndex(d, c)]));
    return k == KIND_SOLID || k == KIND_POWDER;
}
__device__ __forceinline__ bool walls(const Dev &d, i2 c) {               // :1056-1062
    if (!inWorld(d, c) || isOccupant(occupantAt(d, cellIndex(d, c)))) return true;
    uint32_t k = kindOf(d, wType(d.cells[4ull * cellIndex(d, c)]));
    return k == KIND_SOLID || k == KIND_POWDER;
}

// union-find over G_REGION words (label.wgsl:25-55)
__device__ __forceinline__ uint32_t ufLink(const Dev &d, uint32_t at) { return ld(&CNT(d, at, G_REGION)); }
// (with path halving: a cell's link is lowered to its grandparent on the
// way up, an ancestor, so every find still ends at the same root; a root's
// own word, where the region's open bit goes, is never written here)
__device__ uint32_t ufFind(const Dev &d, uint32_t i) {
    uint32_t x = i;
    for (;;) {
        uint32_t up = ufLink(d, x);
        if (up >= x) break;
        uint32_t gp = ufLink(d, up);
        if (gp >= up) {
            x = up;
            break;
        }
        atomicMin(&CNT(d, x, G_REGION), gp);
        x = gp;
    }
    return x;
}
__device__ void ufJoin(const Dev &d, uint32_t i, uint32_t j) {
    uint32_t a = i, b = j;
    for (;;) {
        a = ufFind(d, a);
        b = ufFind(d, b);
        if (a == b) return;
        uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        uint32_t old = atomicMin(&CNT(d, hi, G_REGION), lo);
        if (old == hi) return;
        a = old;
        b = lo;
    }
}
__device__ __forceinline__ bool pinnedAt(const Dev &d, uint32_t i) { return ufLink(d, i) == i; }

// facesOf (:296-303)
struct Face3 { int32_t x, y, isU; };
__device__ __forceinline__ void facesOf(const Dev &d, i2 c, Face3 fs[4]) {
    bool right = c.x == d.W - 1 || (c.x % WIG_BLOCK == WIG_BLOCK - 1 && !isActiveBlock(d, I2(c.x / WIG_BLOCK + 1, c.y / WIG_BLOCK)));
    bool bottom = c.y == d.H - 1 || (c.y % WIG_BLOCK == WIG_BLOCK - 1 && !isActiveBlock(d, I2(c.x / WIG_BLOCK, c.y / WIG_BLOCK + 1)));
    fs[0].x = c.x; fs[0].y = c.y; fs[0].isU = 1;
    fs[1].x = c.x; fs[1].y = c.y; fs[1].isU = 0;
    fs[2].x = c.x + 1; fs[2].y = c.y; fs[2].isU = right ? 1 : -1;
    fs[3].x = c.x; fs[3].y = c.y + 1; fs[3].isU = bottom ? 0 : -1;
}

// the cell of thread li of the block at slot wid of the active list
__device__ __forceinline__ i2 blockCell(i2 b, uint32_t li) { return I2(b.x * WIG_BLOCK + (int32_t)(li % 8u), b.y * WIG_BLOCK + (int32_t)(li / 8u)); }

// (the first block's list entry is read with the count, in one round trip;
// the next as the loop goes: blk_, activeBlockAt)
#define FOR_ACTIVE_BLOCKS(d, wid) \
    for (uint32_t wid = bid, n_ = activeCount(d), blk_ = (d).blocks[ACTIVE_SLOT(d, bid < (d).B ? bid : 0u)]; wid < n_; \
         wid += nbid, blk_ = wid < n_ ? (d).blocks[ACTIVE_SLOT(d, wid)] : 0u)
__device__ __forceinline__ i2 activeBlockAt(const Dev &d, uint32_t b) { return I2((int32_t)(b % d.bw), (int32_t)(b / d.bw)); }

// -------------------------------------------------- wavefront reductions
__device__ __forceinline__ double wave_tree_sum(double v) {
    // the 64 values summed by halving: lane i takes lane i + h, h = 32 .. 1
    for (int h = 32; h > 0; h >>= 1) v += __shfl_down(v, h, WAVE);
    return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int h = 32; h > 0; h >>= 1) { uint32_t o = __shfl_down(v, h, WAVE); v = o > v ? o : v; }
    return v;
}
// inclusive prefix sum over the wavefront
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    uint32_t lane = lane_id();
    for (int o = 1; o < WAVE; o <<= 1) {
        uint32_t t = __shfl_up(v, o, WAVE);
        if (lane >= (uint32_t)o) v += t;
    }
    return v;
}

// ================================================================== kernels

// Zeroes n words (device memory set without hipMemset, so a captured graph
// holds only kernels).
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

// sync / syncMarks (:608-659), a workgroup per block of the world
__global__ void __launch_bounds__(64) k_sync(WorldGrid g) {
    WORLD(g);
    const int claims = d.full ? 1 : 0;
    const uint32_t gx_ = d.bw, bx = bid % gx_, by = bid / gx_;
    uint32_t li = threadIdx.x;
    i2 c = I2((int32_t)bx * WIG_BLOCK + (int32_t)(li % 8u), (int32_t)by * WIG_BLOCK + (int32_t)(li / 8u));
    bool liquid = false, region = false;
    if (inWorld(d, c)) {
        uint32_t i = cellIndex(d, c);
        if (i == 0u) {
            d.parts[0] = 0u;
            d.blocks[COUNT_SLOT(d, COUNT_ACTIVE)] = 0u;
            d.blocks[COUNT_SLOT(d, COUNT_RETIRED)] = 0u;
            d.blocks[COUNT_SLOT(d, COUNT_VISCOUS)] = 0u;
        }
        uint4 v = *(const uint4 *)&d.cells[4ull * i];
        if (kindOf(d, wType(v.x)) == KIND_LIQUID) {
            uint32_t home = v.w >> 8;
            if (claims && wHas(v.x, FL_DRAWN) && home < d.C) atomicMax(&CNT(d, home, G_DEST), DEST_CELL | (DEST_SPAN - i));
            liquid = true;
        } else if (d.body_on) {
            region = wetRegionAt(d, v);
        }
    }
    if (__ballot(liquid) && li == 0) atomicOr(&d.blocks[by * d.bw + bx], MARK_LIQUID);
    if (__ballot(region) && li == 0) atomicOr(&d.blocks[by * d.bw + bx], MARK_REGION);
}

__device__ __forceinline__ void markBlockOf(const Dev &d, v2 pos) {
    i2 b = idiv8(cellOf(d, pos));
    if (inBlocks(d, b)) atomicOr(&d.blocks[blockIndex(d, b)], MARK_LIQUID);
}

__device__ v2 pairWay(uint32_t a, uint32_t b) {                          // :703-707
    uint32_t h = pcg((a < b ? a : b) ^ pcg(a < b ? b : a));
    float angle = (float)h * (float)(6.2831853 / 4294967296.0);
    float s = a < b ? 1.0f : -1.0f;
    return V2(cosf(angle) * s, sinf(angle) * s);
}
__device__ v2 apart(const Dev &d, uint32_t k, uint32_t h, v2 pos, uint32_t id) {   // :713-734
    uint32_t start = CNT(d, h, G_START), n = CNT(d, h, G_SORT);
    if (n < 2u) return pos;
    int32_t px = 0, py = 0;
    for (uint32_t m = start; m < start + n; m++) {
        if (m == k) continue;
        v2 dd = V2(pos.x - fbits(MOTION(d, d.sorted, m, 0)), pos.y - fbits(MOTION(d, d.sorted, m, 1)));
        float l = wlength(dd);
        if (l < APART) {
            v2 way = l == 0.0f ? pairWay(id, STATE(d, d.sorted, m, 2)) : V2(dd.x / l, dd.y / l);
            float s = 0.5f * (APART - l) * APART_FIX;
            px += wi32(wround(s * way.x));
            py += wi32(wround(s * way.y));
        }
    }
    v2 c = V2(floorf(pos.x), floorf(pos.y));
    v2 to = V2(pos.x + (float)px / APART_FIX, pos.y + (float)py / APART_FIX);
    v2 r;
    r.x = (to.x >= c.x + APART && to.x <= c.x + 1.0f - APART) ? to.x : pos.x;
    r.y = (to.y >= c.y + APART && to.y <= c.y + 1.0f - APART) ? to.y : pos.y;
    return r;
}
__device__ uint32_t rankById(const Dev &d, uint32_t h, uint32_t k, uint32_t id, uint32_t word, uint32_t drawn) {   // :668-678
    uint32_t start = CNT(d, h, G_START), mine = wType(word) == drawn ? 1u : 0u, r = 0;
    for (uint32_t m = start; m < start + CNT(d, h, G_SORT); m++) {
        uint32_t ox = STATE(d, d.sorted, m, 0), oz = STATE(d, d.sorted, m, 2);
        uint32_t theirs = wType(ox) == drawn ? 1u : 0u;
        r += (theirs < mine || (theirs == mine && (oz < id || (oz == id && m < k)))) ? 1u : 0u;
    }
    return r;
}
__device__ uint32_t sharedLife(const Dev &d, uint32_t life, uint32_t was, uint32_t now, uint32_t id) {   // :4149-4161
    if (now == was) return life;
    float l = (float)life;
    if (now < was) l = l * (float)now / (float)was;
    else l = 255.0f - (255.0f - l) * (255.0f - (float)now) / (255.0f - (float)was);
    Rng r;
    seedRng(r, id, d.frame, d.seed ^ 0x5a17u);
    uint32_t v = wu32(floorf(l + rnd(r)));
    return v < 255u ? v : 255u;
}

struct PState { v2 pos, vel; uint32_t word; float temp; uint32_t id; float aff[4]; };

__device__ __forceinline__ void storeParticle(const Dev &d, uint32_t m, const PState &q) {   // :534-538
    *(uint4 *)&MOTION(d, d.parts, m, 0) = make_uint4(ubits(q.pos.x), ubits(q.pos.y), ubits(q.vel.x), ubits(q.vel.y));
    *(uint4 *)&STATE(d, d.parts, m, 0) = make_uint4(q.word, ubits(q.temp), q.id, 0u);
    *(uint4 *)&AFFINE(d, d.parts, m, 0) = make_uint4(ubits(q.aff[0]), ubits(q.aff[1]), ubits(q.aff[2]), ubits(q.aff[3]));
}

// Places in `parts` for a wavefront's particles, in lane order: one atomic
// a wavefront (takePlaces, :743-755, takes one a workgroup).
__device__ __forceinline__ uint32_t wave_take(const Dev &d, uint32_t n) {
    uint32_t incl = wave_incl_scan(n);
    uint32_t total = __shfl(incl, WAVE - 1, WAVE);
    uint32_t base = 0;
    if (lane_id() == WAVE - 1 && total) base = atomicAdd(&d.parts[0], total);
    base = __shfl(base, WAVE - 1, WAVE);
    return base + incl - n;
}

// syncParticles (:767-831).  A workgroup takes 1024 consecutive particles
// of `sorted`, a wavefront 256 of them (four rounds of 64, held in
// registers), and puts those it keeps into `parts` in their order with one
// atomic (takePlaces, :743-755, takes one a workgroup of 256 in no set
// order); the kept ones stay together in runs as long as the chunk, which
// keeps the gather's tiles small.
#define SP_CHUNK 1024u
__device__ __forceinline__ bool syncOne(const Dev &d, uint32_t k, PState &q) {
    uint4 mo = *(const uint4 *)&MOTION(d, d.sorted, k, 0);
    uint4 st = *(const uint4 *)&STATE(d, d.sorted, k, 0);
    uint4 af = *(const uint4 *)&AFFINE(d, d.sorted, k, 0);
    q.pos = V2(fbits(mo.x), fbits(mo.y));
    q.vel = V2(fbits(mo.z), fbits(mo.w));
    q.word = st.x;
    q.temp = fbits(st.y);
    q.id = st.z;
    q.aff[0] = fbits(af.x); q.aff[1] = fbits(af.y); q.aff[2] = fbits(af.z); q.aff[3] = fbits(af.w);
    uint32_t h = cellIndex(d, cellOf(d, q.pos));
    q.pos = apart(d, k, h, q.pos, q.id);
    uint32_t dest = CNT(d, h, G_DEST);
    if (dest == DEST_KEEP) {
        markBlockOf(d, q.pos);
        return true;
    }
    if ((dest & DEST_ASIDE) != 0u) {
        uint32_t to = dest & ~DEST_ASIDE;
        // not into another body's cup (:785-794)
        if (cupOf(d, to) == WNONE)
            q.pos = V2((float)(to % (uint32_t)d.W) + (q.pos.x - floorf(q.pos.x)), (float)(to / (uint32_t)d.W) + (q.pos.y - floorf(q.pos.y)));
        markBlockOf(d, q.pos);
        return true;
    }
    if ((dest & DEST_CELL) == 0u) return false;
    uint32_t to = DEST_SPAN - (dest & DEST_SPAN);
    uint4 v = *(const uint4 *)&d.cells[4ull * to];
    uint32_t cn = v.w & LIQUID_COUNT_MAX;
    if (!(cn == LIQUID_COUNT_MAX || CNT(d, h, G_SORT) <= cn || rankById(d, h, k, q.id, q.word, wType(v.x)) < cn)) return false;
    if (to != h)
        q.pos = V2((float)(to % (uint32_t)d.W) + (q.pos.x - floorf(q.pos.x)), (float)(to / (uint32_t)d.W) + (q.pos.y - floorf(q.pos.y)));
    q.temp = fbits(v.y);
    uint32_t drawn = CNT(d, h, G_TYPE), pt = wType(q.word);
    if (pt == E_WATER && pt == (drawn & 0xffu) && wType(v.x) == pt
        && (wAux(q.word) == ((drawn >> 8) & 0xffu) || wAux(q.word) == 0u)) {
        uint32_t life = sharedLife(d, wAux(q.word) == 0u ? 0u : wLife(q.word), (drawn >> 16) & 0xffu, wLife(v.x), q.id);
        uint32_t ww = wWithout(v.x, FL_DRAWN);
        uint32_t solute = wAux(ww) != 0u ? wAux(ww) : (drawn >> 8) & 0xffu;
        q.word = packWord(pt, life > 0u ? solute : 0u, life, wShade(ww), wFlags(ww));
    } else if ((wType(q.word) | (wAux(q.word) << 8)) == (drawn & 0xffffu)) {
        q.word = wWithout(v.x, FL_DRAWN);
    }
    return true;
}
__global__ void __launch_bounds__(256) k_syncParticles(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    __shared__ uint32_t sKept[WAVES + 1];
    uint32_t cnt = ld(&d.sorted[0]);
    uint32_t n = cnt < d.cap ? cnt : d.cap;
    const uint32_t lane = lane_id(), wave = threadIdx.x / 64u;
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t base = bid * SP_CHUNK; base < n; base += nbid * SP_CHUNK) {
        PState q[4];
        bool keep[4];
        uint32_t before[4], total = 0;
#pragma unroll
        for (uint32_t r = 0; r < 4u; r++) {
            const uint32_t k = base + 256u * wave + 64u * r + lane;
            keep[r] = k < n && syncOne(d, k, q[r]);
            const uint64_t kept = __ballot(keep[r]);
            before[r] = total + (uint32_t)__popcll(kept & below);
            total += (uint32_t)__popcll(kept);
        }
        if (lane == 0u) sKept[wave] = total;
        __syncthreads();
        if (threadIdx.x == 0u) {
            uint32_t all = sKept[0] + sKept[1] + sKept[2] + sKept[3];
            sKept[WAVES] = all ? atomicAdd(&d.parts[0], all) : 0u;
        }
        __syncthreads();
        uint32_t at = sKept[WAVES];
        for (uint32_t w = 0; w < wave; w++) at += sKept[w];
#pragma unroll
        for (uint32_t r = 0; r < 4u; r++) {
            const uint32_t m = at + before[r];
            if (keep[r] && m < d.cap) storeParticle(d, m, q[r]);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ uint32_t seedId(const Dev &d, i2 c, uint32_t j) {   // :835-841
    uint32_t h = (((uint32_t)c.y * (uint32_t)d.W + (uint32_t)c.x) * 0x9e3779b9u) ^ (d.frame * 0x85ebca6bu) ^ (j * 0xc2b2ae35u);
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    return h;
}

// syncSeed (:863-890), a workgroup per block of the world
__global__ void __launch_bounds__(64) k_syncSeed(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    const uint32_t gx_ = d.bw, bx = bid % gx_, by = bid / gx_;
    uint32_t li = threadIdx.x;
    i2 c = I2((int32_t)bx * WIG_BLOCK + (int32_t)(li % 8u), (int32_t)by * WIG_BLOCK + (int32_t)(li / 8u));
    uint32_t n = 0;
    uint4 v = make_uint4(0, 0, 0, 0);
    Rng r;
    r.s = 0;
    uint32_t i = 0;
    if (inWorld(d, c)) {
        i = cellIndex(d, c);
        v = *(const uint4 *)&d.cells[4ull * i];
        if (kindOf(d, wType(v.x)) == KIND_LIQUID) {
            uint32_t home = v.w >> 8, want = v.w & LIQUID_COUNT_MAX;
            if (!wHas(v.x, FL_DRAWN)) n = LIQUID_PER_CELL;
            else if (home >= d.C || CNT(d, home, G_DEST) != (DEST_CELL | (DEST_SPAN - i))) n = want > 1u ? want : 1u;
            else if (want > CNT(d, home, G_SORT) && want < LIQUID_COUNT_MAX) n = want - CNT(d, home, G_SORT);
            seedRng(r, i, d.frame, d.seed ^ 0x5eedu);
        }
    }
    uint32_t m = wave_take(d, n);
    for (uint32_t j = 0; j < n; j++) {
        PState q;
        if (n == 1u) {
            q.pos = V2((float)c.x + 0.5f, (float)c.y + 0.5f);
        } else {
            float rx = rnd(r), ry = rnd(r);
            v2 rr = V2(rx - 0.5f, ry - 0.5f);
            if (n <= 4u) {
                float qx = (j == 1u || j == 2u) ? 1.0f : 0.0f, qy = (j == 1u || j == 3u) ? 1.0f : 0.0f;
                q.pos = V2((float)c.x + 0.25f + 0.5f * qx + 0.3f * rr.x, (float)c.y + 0.25f + 0.5f * qy + 0.3f * rr.y);
            } else {
                q.pos = V2((float)c.x + 0.5f + 0.9f * rr.x, (float)c.y + 0.5f + 0.9f * rr.y);
            }
        }
        q.vel = unpack2x16(v.z);
        q.word = wWithout(v.x, FL_DRAWN);
        q.temp = fbits(v.y);
        q.id = seedId(d, c, j);
        q.aff[0] = q.aff[1] = q.aff[2] = q.aff[3] = 0.0f;
        if (m + j < d.cap) storeParticle(d, m + j, q);
    }
}

// compact (:896-921): the blocks worked on and those dropping out, listed in
// block-index order by one workgroup (the reference's lists are in whatever
// order its threads came).
// compact's per-block flags (:896-921): each block's 3 x 3 neighbourhood of
// marks, one thread a block
__global__ void __launch_bounds__(256) k_compactMarks(WorldGrid g) {
    WORLD(g);
    uint32_t n = bid * 256u + threadIdx.x;
    if (n >= d.B) return;
    i2 b = I2((int32_t)(n % d.bw), (int32_t)(n / d.bw));
    uint32_t marks = 0;
    for (int y = -1; y <= 1; y++)
        for (int x = -1; x <= 1; x++) {
            i2 o = I2(b.x + x, b.y + y);
            if (inBlocks(d, o)) marks |= d.blocks[blockIndex(d, o)];
        }
    bool nearl = (marks & MARK_LIQUID) != 0u, liquid = (marks & (MARK_LIQUID | MARK_REGION)) != 0u;
    bool was = (d.blocks[STATE_SLOT(d, n)] & ACTIVE) != 0u;
    d.scratch[n] = (liquid ? 1u : 0u) | (!liquid && was ? 2u : 0u) | (nearl ? 4u : 0u);
}
// compact's lists, in block-index order, by one workgroup of 1024: a chunk
// of 16384 blocks at a time, each wave taking 1024 consecutive blocks in 16
// rows of 64 (lane l of row q takes block 64q + l of its wave's), so that
// the flags' loads, the block states' stores and the lists' stores are all
// coalesced.  A block's place in a list is the blocks listed before its row
// (earlier chunks, earlier waves, earlier rows) plus those before it in its
// row (a ballot's bits below its lane).
#define COMPACT_ROWS 16
#define COMPACT_CHUNK (COMPACT_ROWS * 1024u)
__global__ void __launch_bounds__(1024) k_compact(WorldGrid g) {
    WORLD(g);
    __shared__ uint32_t sA[16], sR[16];
    const uint32_t lane = lane_id(), wave = threadIdx.x / 64u;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t baseA = 0, baseR = 0;
    for (uint32_t c0 = 0; c0 < d.B; c0 += COMPACT_CHUNK) {
        const uint32_t w0 = c0 + wave * (COMPACT_ROWS * 64u);
        uint32_t f[COMPACT_ROWS];
#pragma unroll
        for (int q = 0; q < COMPACT_ROWS; q++) {
            const uint32_t n = w0 + 64u * (uint32_t)q + lane;
            f[q] = n < d.B ? d.scratch[n] : 0u;
        }
        uint32_t na = 0, nr = 0;
#pragma unroll
        for (int q = 0; q < COMPACT_ROWS; q++) {
            const uint32_t n = w0 + 64u * (uint32_t)q + lane;
            if (n < d.B) d.blocks[STATE_SLOT(d, n)] = ((f[q] & 1u) ? ACTIVE : 0u) | ((f[q] & 4u) ? NEAR_LIQUID : 0u);
            na += (uint32_t)__popcll(__ballot((f[q] & 1u) != 0u));
            nr += (uint32_t)__popcll(__ballot((f[q] & 2u) != 0u));
        }
        if (lane == 0u) {
            sA[wave] = na;
            sR[wave] = nr;
        }
        __syncthreads();
        uint32_t ia = baseA, ir = baseR;
        for (uint32_t w = 0; w < wave; w++) {
            ia += sA[w];
            ir += sR[w];
        }
        for (uint32_t w = 0; w < 16u; w++) {
            baseA += sA[w];
            baseR += sR[w];
        }
#pragma unroll
        for (int q = 0; q < COMPACT_ROWS; q++) {
            const uint32_t n = w0 + 64u * (uint32_t)q + lane;
            const uint64_t ma = __ballot((f[q] & 1u) != 0u), mr = __ballot((f[q] & 2u) != 0u);
            if (f[q] & 1u) d.blocks[ACTIVE_SLOT(d, ia + (uint32_t)__popcll(ma & below))] = n;
            if (f[q] & 2u) d.blocks[RETIRED_SLOT(d, ir + (uint32_t)__popcll(mr & below))] = n;
            ia += (uint32_t)__popcll(ma);
            ir += (uint32_t)__popcll(mr);
        }
        __syncthreads();   // every wave has read sA, sR before the next chunk's
    }
    if (threadIdx.x == 0u) {
        d.blocks[COUNT_SLOT(d, COUNT_ACTIVE)] = baseA;
        d.blocks[COUNT_SLOT(d, COUNT_RETIRED)] = baseR;
    }
}

// retire (:924-951)
__global__ void __launch_bounds__(64) k_retire(WorldGrid g) {
    WORLD(g);
    for (uint32_t wid = bid, nr = d.blocks[COUNT_SLOT(d, COUNT_RETIRED)]; wid < nr; wid += nbid) {
        uint32_t n = d.blocks[RETIRED_SLOT(d, wid)];
        i2 c = blockCell(I2((int32_t)(n % d.bw), (int32_t)(n / d.bw)), threadIdx.x);
        if (inWorld(d, c)) {
            uint32_t i = cellIndex(d, c);
            *(float4 *)&FLD(d, 2u * i, 0) = make_float4(0.f, 0.f, 0.f, 0.f);
            *(float4 *)&FLD(d, 2u * i + 1u, 0) = make_float4(0.f, 0.f, 0.f, 0.f);
            for (uint32_t g = 0; g < LIQUID_CELL_WORDS; g++) CNT(d, i, g) = 0u;
            Face3 fs[4];
            facesOf(d, c, fs);
            for (int j = 0; j < 4; j++) {
                if (fs[j].isU < 0) continue;
                uint32_t f = faceIndex(d, I2(fs[j].x, fs[j].y), fs[j].isU == 1);
                *(float4 *)&FLD(d, faceSlot(d, f), 0) = make_float4(0.f, 0.f, 0.f, 0.f);
                FLD(d, viscLastSlot(d, f), f % 4u) = 0.0f;
                SUM(d, f, 0) = 0u;
                SUM(d, f, 1) = 0u;
            }
        }
        if (threadIdx.x == 0) d.blocks[n] = 0u;
    }
}

// clear (:962-985)
__global__ void __launch_bounds__(64) k_clear(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 b = activeBlockAt(d, blk_);
        if (threadIdx.x == 0) d.blocks[blockIndex(d, b)] = 0u;
        i2 c = blockCell(b, threadIdx.x);
        if (!inWorld(d, c)) continue;
        uint32_t i = cellIndex(d, c);
        for (uint32_t g = 0; g < LIQUID_CELL_WORDS; g++) CNT(d, i, g) = 0u;
        CNT(d, i, G_REGION) = i;
        Face3 fs[4];
        facesOf(d, c, fs);
        for (int j = 0; j < 4; j++) {
            if (fs[j].isU < 0) continue;
            uint32_t f = faceIndex(d, I2(fs[j].x, fs[j].y), fs[j].isU == 1);
            SUM(d, f, 0) = 0u;
            SUM(d, f, 1) = 0u;
        }
    }
}

// gather (:1000-1087): the splat as fixed-point integers, whose sums are the
// same in any order.  A workgroup takes 256 consecutive particles (which the
// sort by cell keeps close together) and sums their splats into a tile in
// LDS covering their cells and faces, then adds the tile's nonzero words
// into the counters: a few global atomics a cell rather than 22 a particle
// on the most contended words.  Particles too spread for a tile add
// straight into the counters.
#define GT_MAX 1024   // tile entries (cells, and faces of each axis)
struct GatherTile {
    int32_t ox, oy, tw, th;     // origin and size, in cells
    uint32_t *um, *uw, *vm, *vw, *cnt, *mass, *dens;
    __device__ __forceinline__ int at(i2 c) const {
        int32_t x = c.x - ox, y = c.y - oy;
        return (x >= 0 && y >= 0 && x < tw && y < th) ? y * tw + x : -1;
    }
};
template <bool TILE>
__device__ __forceinline__ void addWord(uint32_t *tile, int t, uint32_t *global, uint32_t v) {
    if (TILE && t >= 0) ldsAdd(&tile[t], v);
    else glbAdd(global, v);
}
template <bool TILE>
__device__ __forceinline__ void splatAxis(const Dev &d, const GatherTile &T, v2 s, float vel, float gx, float gy, bool isU) {
    float fx0 = floorf(s.x), fy0 = floorf(s.y);
    i2 b = I2(wi32(fx0), wi32(fy0));
    v2 f = V2(s.x - fx0, s.y - fy0);
    for (int j = 0; j < 4; j++) {
        int ox = j & 1, oy = j >> 1;
        float wgt = (ox == 1 ? f.x : 1.0f - f.x) * (oy == 1 ? f.y : 1.0f - f.y);
        i2 at = I2(b.x + ox, b.y + oy);
        if (wgt <= 0.0f || !faceInWorld(d, at, isU)) continue;
        uint32_t fi = faceIndex(d, at, isU);
        float there = vel + (gx * ((float)at.x - s.x) + gy * ((float)at.y - s.y));
        uint32_t m = (uint32_t)wi32(wround(there * wgt * FIXF)), w = (uint32_t)wi32(wround(wgt * FIXF));
        int t = TILE ? T.at(at) : -1;
        addWord<TILE>(isU ? T.um : T.vm, t, &SUM(d, fi, 0), m);
        addWord<TILE>(isU ? T.uw : T.vw, t, &SUM(d, fi, 1), w);
    }
}
// splatDensity's four cells, b .. b + 1 from its corner b = floor(pos - 0.5),
// and their first words (0 off the world), loaded when the particle's are
struct DensityCells { i2 b; uint32_t w[4], occ[4]; };
__device__ __forceinline__ DensityCells densityCells(const Dev &d, v2 pos) {
    DensityCells q;
    q.b = I2(wi32(floorf(pos.x - 0.5f)), wi32(floorf(pos.y - 0.5f)));
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const i2 c = I2(q.b.x + (j & 1), q.b.y + (j >> 1));
        const uint32_t ci = cellIndex(d, clampTo(c, d.W, d.H));
        const uint32_t w = d.cells[4ull * ci], o = occupantAt(d, ci);
        q.w[j] = inWorld(d, c) ? w : 0u;
        q.occ[j] = inWorld(d, c) ? o : 0u;
    }
    return q;
}
template <bool TILE>
__device__ __forceinline__ void splatDensity(const Dev &d, const GatherTile &T, v2 pos, float room, const DensityCells &q,
                                             const uint8_t *kinds) {
    v2 s = V2(pos.x - 0.5f, pos.y - 0.5f);
    float fx0 = floorf(s.x), fy0 = floorf(s.y);
    i2 b = I2(wi32(fx0), wi32(fy0));
    v2 f = V2(s.x - fx0, s.y - fy0);
    // walls() (:1056-1062) of the four cells: off the world, someone's
    // body, solid or powder
    bool wall[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t k = ldsGetU8(&kinds[wType(q.w[j]) & 0xffu]);
        wall[j] = !inWorld(d, I2(b.x + (j & 1), b.y + (j >> 1))) || isOccupant(q.occ[j]) || k == KIND_SOLID || k == KIND_POWDER;
    }
    for (int j = 0; j < 4; j++) {
        int ox = j & 1, oy = j >> 1;
        float wgt = (ox == 1 ? f.x : 1.0f - f.x) * (oy == 1 ? f.y : 1.0f - f.y);
        i2 at = I2(b.x + ox, b.y + oy);
        if (wall[j]) {
            // the corner across in x, in y, then the diagonal
            i2 x = I2(b.x + 1 - ox, at.y), y = I2(at.x, b.y + 1 - oy), xy = I2(b.x + 1 - ox, b.y + 1 - oy);
            at = !wall[j ^ 1] ? x : (!wall[j ^ 2] ? y : (!wall[j ^ 3] ? xy : I2(-1, -1)));
        }
        if (inWorld(d, at) && wgt > 0.0f)
            addWord<TILE>(T.dens, TILE ? T.at(at) : -1, &CNT(d, cellIndex(d, at), G_DENSITY), wu32(wround(wgt * room * FIXF)));
    }
}
// (particle k's motion, state and affine words and its density cells,
// loaded by the caller; kinds: the element kinds staged in LDS)
template <bool TILE>
__device__ __forceinline__ void splatParticle(const Dev &d, const GatherTile &T, uint32_t k, uint4 mo, const uint4 st, const uint4 af,
                                              const DensityCells &q, const uint8_t *kinds) {
    v2 pos = V2(fbits(mo.x), fbits(mo.y));
    // the air's pressure on a particle beside gas (:1024-1032): its
    // velocity pushed, and kept
    if (d.air_on) {
        const i2 c = I2(wi32(floorf(pos.x)), wi32(floorf(pos.y)));
        bool gas = false;
        for (int q = 0; q < 4; q++) {
            const i2 n = beside(c, q);
            gas = gas || (inWorld(d, n) && kindOf(d, wType(d.cells[4ull * cellIndex(d, n)])) == KIND_GAS);
        }
        if (gas) {
            const v2 push = pressurePush(d, airGrad(d, pos), WATER_RHO);
            mo.z = ubits(fbits(mo.z) + push.x);
            mo.w = ubits(fbits(mo.w) + push.y);
            MOTION(d, d.parts, k, 2) = mo.z;
            MOTION(d, d.parts, k, 3) = mo.w;
        }
    }
    float temp = fbits(st.y);
    splatAxis<TILE>(d, T, V2(pos.x - 0.0f, pos.y - 0.5f), fbits(mo.z), fbits(af.x), fbits(af.y), true);
    splatAxis<TILE>(d, T, V2(pos.x - 0.5f, pos.y - 0.0f), fbits(mo.w), fbits(af.z), fbits(af.w), false);
    i2 hc = cellOf(d, pos);
    uint32_t home = cellIndex(d, hc);
    int t = TILE ? T.at(hc) : -1;
    addWord<TILE>(T.cnt, t, &CNT(d, home, G_COUNT), 1u);
    addWord<TILE>(T.mass, t, &CNT(d, home, G_MASS), wu32(wround(liquidDensity(temp) * FIXF)));
    splatDensity<TILE>(d, T, pos, liquidVolume(temp), q, kinds);
}
// (the tile's words, added into the counters: one atomic a nonzero word;
// threads t0, t0 + step, ... of the tile's entries)
__device__ __forceinline__ void flushTile(const Dev &d, const GatherTile &T, int t0, int step) {
    const int ne = T.tw * T.th;
    for (int j = t0; j < ne; j += step) {
        i2 c = I2(T.ox + j % T.tw, T.oy + j / T.tw);
        const uint32_t um = ldsGet(&T.um[j]), uw = ldsGet(&T.uw[j]), vm = ldsGet(&T.vm[j]), vw = ldsGet(&T.vw[j]);
        const uint32_t cn = ldsGet(&T.cnt[j]), ms = ldsGet(&T.mass[j]), dn = ldsGet(&T.dens[j]);
        if (um | uw) {          // (a face outside the world was never added to)
            uint32_t fi = uFace(d, c.x, c.y);
            if (um) glbAdd(&SUM(d, fi, 0), um);
            if (uw) glbAdd(&SUM(d, fi, 1), uw);
        }
        if (vm | vw) {
            uint32_t fi = vFace(d, c.x, c.y);
            if (vm) glbAdd(&SUM(d, fi, 0), vm);
            if (vw) glbAdd(&SUM(d, fi, 1), vw);
        }
        if (cn | ms | dn) {
            uint32_t ci = cellIndex(d, c);
            if (cn) glbAdd(&CNT(d, ci, G_COUNT), cn);
            if (ms) glbAdd(&CNT(d, ci, G_MASS), ms);
            if (dn) glbAdd(&CNT(d, ci, G_DENSITY), dn);
        }
    }
}
__device__ __forceinline__ void tileWords(GatherTile &T, uint32_t *base) {
    const int ne = T.tw * T.th;
    T.um = base;
    T.uw = base + ne;
    T.vm = base + 2 * ne;
    T.vw = base + 3 * ne;
    T.cnt = base + 4 * ne;
    T.mass = base + 5 * ne;
    T.dens = base + 6 * ne;
}
// A workgroup takes 256 particles at a time.  Their cells' box (the faces
// and cells they touch) in LDS when it fits; when it does not (the 256 span
// two places, as runs of particles from different regions meet), each
// wavefront's own box in a quarter of it; past that, the counters directly.
__global__ void __launch_bounds__(256) k_gather(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    __shared__ uint32_t sTile[7 * GT_MAX];
    __shared__ int32_t sBox[4];
    uint32_t cnt = ld(&d.parts[0]);
    uint32_t n = cnt < d.cap ? cnt : d.cap;
    __shared__ uint8_t sKind[256];
    const uint32_t tid = threadIdx.x, wave = tid / 64u;
    sKind[tid] = (uint8_t)d.kind[tid];   // (the first box's barriers pass before a read)
    for (uint32_t base = bid * 256u; base < n; base += nbid * 256u) {
        uint32_t k = base + tid;
        // the particle's words and its density cells' words, all loaded at
        // once, before the box is known
        const uint32_t kk = k < n ? k : n - 1u;
        const uint4 mo = *(const uint4 *)&MOTION(d, d.parts, kk, 0);
        const uint4 st = *(const uint4 *)&STATE(d, d.parts, kk, 0);
        const uint4 af = *(const uint4 *)&AFFINE(d, d.parts, kk, 0);
        const DensityCells q = densityCells(d, V2(fbits(mo.x), fbits(mo.y)));
        // the cells' box, the wavefront's and the workgroup's
        int32_t cx = 0x7fffffff, cy = 0x7fffffff, mx = -1, my = -1;
        if (k < n) {
            i2 c = cellOf(d, V2(fbits(mo.x), fbits(mo.y)));
            cx = mx = c.x;
            cy = my = c.y;
        }
        for (int h = 32; h > 0; h >>= 1) {
            cx = min(cx, __shfl_xor(cx, h, WAVE));
            cy = min(cy, __shfl_xor(cy, h, WAVE));
            mx = max(mx, __shfl_xor(mx, h, WAVE));
            my = max(my, __shfl_xor(my, h, WAVE));
        }
        if (tid == 0) { sBox[0] = 0x7fffffff; sBox[1] = 0x7fffffff; sBox[2] = -1; sBox[3] = -1; }
        __syncthreads();
        if (lane_id() == 0) {
            atomicMin(&sBox[0], cx);
            atomicMin(&sBox[1], cy);
            atomicMax(&sBox[2], mx);
            atomicMax(&sBox[3], my);
        }
        __syncthreads();
        GatherTile T;
        T.ox = sBox[0] - 1;
        T.oy = sBox[1] - 1;
        T.tw = sBox[2] - sBox[0] + 4;   // faces and cells a particle touches: its cell -1 .. +2
        T.th = sBox[3] - sBox[1] + 4;
        const bool whole = T.tw * T.th <= GT_MAX;
        __syncthreads();
        if (whole) {
            tileWords(T, sTile);
            for (int j = (int)tid; j < 7 * T.tw * T.th; j += 256) sTile[j] = 0u;
            __syncthreads();
            if (k < n) splatParticle<true>(d, T, k, mo, st, af, q, sKind);
            __syncthreads();
            flushTile(d, T, (int)tid, 256);
            __syncthreads();
            continue;
        }
        // each wavefront's box (an empty wavefront's is empty)
        GatherTile V;
        V.ox = cx - 1;
        V.oy = cy - 1;
        V.tw = mx >= 0 ? mx - cx + 4 : 0;
        V.th = mx >= 0 ? my - cy + 4 : 0;
        const bool quarter = V.tw * V.th <= GT_MAX / WAVES;
        uint32_t *mine = sTile + wave * (7u * GT_MAX / WAVES);
        tileWords(V, mine);
        if (quarter)
            for (int j = (int)lane_id(); j < 7 * V.tw * V.th; j += 64) mine[j] = 0u;
        __syncthreads();
        if (k < n) {
            if (quarter) splatParticle<true>(d, V, k, mo, st, af, q, sKind);
            else splatParticle<false>(d, V, k, mo, st, af, q, sKind);
        }
        __syncthreads();
        if (quarter) flushTile(d, V, (int)lane_id(), 64);
        __syncthreads();
    }
}

// sealJoin, sealAnchor (:1503-1554)
__device__ __forceinline__ bool mayBeSealed(const Dev &d, i2 c) {
    if (countAt(d, c) != 0) return false;
    uint32_t i = cellIndex(d, c);
    return FLD(d, 2u * i, 2) > 0.0f || FLD(d, 2u * i + 1u, 2) != 0.0f;
}
__device__ __forceinline__ bool wasOpen(const Dev &d, i2 q) {
    if (!inWorld(d, q) || countAt(d, q) < 0) return false;
    return (wu32(FLD(d, 2u * cellIndex(d, q) + 1u, 3)) & stepBit(d, 1u)) == 0u;
}
// The union-find's joins (label.wgsl's), of member cells with their
// member neighbours to the right and below, in two passes: first each
// block's own, in LDS, each member's link set to its component's lowest
// cell in the block; then across the blocks' right and bottom edges, in the
// counters' links.  The roots come out as the joins one at a time would
// leave them: each component's lowest index.
__device__ __forceinline__ uint32_t ldsFind(uint32_t *link, uint32_t i) {
    uint32_t x = i;
    for (;;) {
        uint32_t up = link[x];
        if (up >= x) return x;
        x = up;
    }
}
__device__ void ldsJoin(uint32_t *link, uint32_t i, uint32_t j) {
    uint32_t a = i, b = j;
    for (;;) {
        a = ldsFind(link, a);
        b = ldsFind(link, b);
        if (a == b) return;
        uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        uint32_t old = atomicMin(&link[hi], lo);
        if (old == hi) return;
        a = old;
        b = lo;
    }
}
// member cells: 0 the gaps the liquid left (mayBeSealed), 1 the liquid
__device__ __forceinline__ bool isMember(const Dev &d, i2 c, int mode) {
    if (!inWorld(d, c)) return false;
    return mode == 0 ? mayBeSealed(d, c) : FLD(d, 2u * cellIndex(d, c), 2) > 0.0f;
}
__global__ void __launch_bounds__(64) k_localJoin(WorldGrid g, int mode) {
    WORLD(g);
    if (!d.full) return;
    __shared__ uint32_t sLink[64];
    __shared__ uint32_t sMember[64];
    uint32_t li = threadIdx.x;
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 c = blockCell(activeBlockAt(d, blk_), li);
        bool m = isMember(d, c, mode);
        sLink[li] = li;
        sMember[li] = m ? 1u : 0u;
        __syncthreads();
        if (m && li % 8u < 7u && sMember[li + 1u]) ldsJoin(sLink, li, li + 1u);
        if (m && li / 8u < 7u && sMember[li + 8u]) ldsJoin(sLink, li, li + 8u);
        __syncthreads();
        if (m) {
            uint32_t r = ldsFind(sLink, li);
            i2 rc = blockCell(activeBlockAt(d, blk_), r);   // within a block, index order is li order
            CNT(d, cellIndex(d, c), G_REGION) = cellIndex(d, rc);
        }
        __syncthreads();
    }
}
// (Each join starts from the two cells' links as the local joins left them,
// their blocks' roots; a lane whose pair is its edge neighbour's, the lane
// before it along the same block edge, leaves the join to that lane: the
// same join, and a union done twice is done once.)
__global__ void __launch_bounds__(64) k_edgeJoin(WorldGrid g, int mode) {
    WORLD(g);
    if (!d.full) return;
    const uint32_t li = threadIdx.x;
    const bool right = li % 8u == 7u, down = li / 8u == 7u;
    FOR_ACTIVE_BLOCKS(d, wid) {
        const i2 c = blockCell(activeBlockAt(d, blk_), li);
        const bool mine = (right || down) && isMember(d, c, mode);
        const bool jr = mine && right && isMember(d, I2(c.x + 1, c.y), mode);
        const bool jd = mine && down && isMember(d, I2(c.x, c.y + 1), mode);
        uint32_t a = 0, br = 0, bd = 0;
        if (mine) a = ufLink(d, cellIndex(d, c));
        if (jr) br = ufLink(d, cellIndex(d, I2(c.x + 1, c.y)));
        if (jd) bd = ufLink(d, cellIndex(d, I2(c.x, c.y + 1)));
        // the pair of the lane before along each edge: above on the right
        // edge, to the left on the lower edge
        const uint32_t ar = __shfl(a, (int)((li + 56u) & 63u), WAVE), brr = __shfl(br, (int)((li + 56u) & 63u), WAVE);
        const bool jrr = __shfl(jr ? 1 : 0, (int)((li + 56u) & 63u), WAVE) != 0;
        const uint32_t ad = __shfl(a, (int)((li + 63u) & 63u), WAVE), bdd = __shfl(bd, (int)((li + 63u) & 63u), WAVE);
        const bool jdd = __shfl(jd ? 1 : 0, (int)((li + 63u) & 63u), WAVE) != 0;
        if (jr && !(li >= 8u && jrr && ar == a && brr == br)) ufJoin(d, a, br);
        if (jd && !(li % 8u > 0u && jdd && ad == a && bdd == bd)) ufJoin(d, a, bd);
    }
}
// sealAnchor: a gap meeting what was open air last step opens its region
__global__ void __launch_bounds__(64) k_sealAnchor(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (!inWorld(d, c) || !mayBeSealed(d, c)) continue;
        if (wasOpen(d, beside(c, 0)) || wasOpen(d, beside(c, 1)) || wasOpen(d, beside(c, 2)) || wasOpen(d, beside(c, 3))) {
            uint32_t r = ufFind(d, cellIndex(d, c));
            if ((ufLink(d, r) & REGION_OPEN) == 0u) atomicOr(&CNT(d, r, G_REGION), REGION_OPEN);   // (one already set needs none)
        }
    }
}

// classify (:1566-1696)
__device__ float columnWeight(const Dev &d, i2 p, float density) {
    int over = 0;
    i2 q = I2(p.x, p.y - 1);
    while (over < COLUMN_MOST && countAt(d, q) > 0) {
        over += 1;
        q.y -= 1;
    }
    return density * d.earthG * ((float)over + 0.5f);
}
__global__ void __launch_bounds__(64) k_classify(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 bk = activeBlockAt(d, blk_);
        i2 p = blockCell(bk, threadIdx.x);
        uint32_t bits = 0;
        if (inWorld(d, p)) {
            uint32_t i = cellIndex(d, p);
            int32_t n = countAt(d, p);
            float4 was = *(const float4 *)&FLD(d, 2u * i, 0);
            float4 wasDrift = *(const float4 *)&FLD(d, 2u * i + 1u, 0);
            float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            bool sealed = false;
            if (n == 0 && mayBeSealed(d, p)) sealed = (ufLink(d, ufFind(d, i)) & REGION_OPEN) == 0u;
            uint32_t rigid = WNONE, held = WNONE;
            const uint32_t seam = n < 0 ? WNONE : seamOf(d, i);
            if (seam != WNONE) {
                // a gap the stamp opened in a body's wall: the body's (:1603-1612)
                c[2] = -1.0f;
                rigid = seam;
                if (isHeld(d, seam)) {
                    held = seam;
                    rigid = WNONE;
                }
                sealed = false;
            } else if (n < 0) {
                c[2] = -1.0f;
                if (d.body_on) {
                    const uint4 v = *(const uint4 *)&d.cells[4ull * i];
                    rigid = rigidBodyOf(d, v);
                    // powder in its cup is its cargo, moving with it (:1616-1619)
                    if (rigid == WNONE && kindOf(d, wType(v.x)) == KIND_POWDER) rigid = cupOf(d, i);
                    if (rigid != WNONE && isHeld(d, rigid)) {
                        held = rigid;
                        rigid = WNONE;
                    }
                }
            } else if (n > 0) {
                v2 li = liquidIn(d, i);
                c[2] = li.x;
                c[3] = li.y;
                // liquid at the surface, where surface tension acts, takes on
                // its viscosity (SURFACE_DAMP, :1626-1633)
                const float sigma = waterTension(d);
                if (sigma > 0.0f && (countAt(d, beside(p, 0)) == 0 || countAt(d, beside(p, 1)) == 0
                                     || countAt(d, beside(p, 2)) == 0 || countAt(d, beside(p, 3)) == 0))
                    c[3] = fmaxf(c[3], sigma * d.eprops[wType(d.cells[4ull * i])].tension * SURFACE_DAMP);
            } else {
                float fluid = 0.0f, nwalls = 0.0f;
                v2 mean = V2(0.0f, 0.0f);
                for (int k = 0; k < 4; k++) {
                    i2 q = beside(p, k);
                    int32_t m = countAt(d, q);
                    if (m < 0) nwalls += 1.0f;
                    else if (m > 0) {
                        fluid += 1.0f;
                        v2 li = liquidIn(d, cellIndex(d, q));
                        mean.x += li.x;
                        mean.y += li.y;
                    }
                }
                if (fluid >= 2.0f && fluid + nwalls >= 3.0f) {
                    c[2] = mean.x / fluid;
                    c[3] = mean.y / fluid;
                } else if (sealed) {
                    c[2] = fluid > 0.0f ? mean.x / fluid : 1.0f;
                    c[3] = fluid > 0.0f ? mean.y / fluid : 0.0f;
                }
            }
            uint32_t mine = stepBit(d, 0u);
            uint32_t liquid = (c[2] > 0.0f || rigid != WNONE || held != WNONE) ? mine : 0u;
            float4 drift = make_float4(0.0f, 0.0f, sealed ? 1.0f : 0.0f, (float)((wu32(wasDrift.w) & ~(mine | fallingBit(d, 0u))) | liquid));
            if (rigid != WNONE) drift.z = -1.0f - (float)rigid;
            else if (held != WNONE) drift.z = -1.0f - (float)(held + MAX_BODIES);
            if (c[2] > 0.0f) {
                if (was.z > 0.0f) {
                    c[0] = was.x;
                    drift.x = wasDrift.x * DRIFT_KEEP;
                } else {
                    c[0] = columnWeight(d, p, c[2]);
                }
                bits = c[3] > 0.0f ? (HAS_FLUID | HAS_VISCOUS) : HAS_FLUID;
            }
            *(float4 *)&FLD(d, 2u * i, 0) = make_float4(c[0], c[1], c[2], c[3]);
            *(float4 *)&FLD(d, 2u * i + 1u, 0) = drift;
        }
        // the block's bits, OR-reduced over the wavefront
        for (int h = 32; h > 0; h >>= 1) bits |= __shfl_xor(bits, h, WAVE);
        if (threadIdx.x == 0) {
            atomicOr(&d.blocks[STATE_SLOT(d, blockIndex(d, bk))], bits);
            if (bits & HAS_VISCOUS) d.blocks[COUNT_SLOT(d, COUNT_VISCOUS)] = 1u;
        }
    }
}

// surface (:1760-1807): on each open cell beside the liquid, σκ, κ from the
// liquid's outline smoothed over 5 x 5 cells, and the contact lines' pull
__device__ __forceinline__ float outline(const Dev &d, i2 c) {            // :1744-1754
    if (!inWorld(d, c)) return 0.5f;
    const uint32_t i = cellIndex(d, c);
    if (isObstacleCell(d, c, d.cells[4ull * i])) return 0.5f;
    return CNT(d, i, G_COUNT) > 0u ? 1.0f : 0.0f;
}
__global__ void __launch_bounds__(64) k_surface(WorldGrid g) {
    WORLD(g);
    if (!d.full || !d.surf_on) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        const i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (!inWorld(d, c)) continue;
        const uint32_t i = cellIndex(d, c);
        if (FLD(d, 2u * i, 2) != 0.0f) continue;
        float tension = 0.0f, contact = 0.0f;
        for (int k = 0; k < 4; k++) {
            const i2 q = beside(c, k);
            if (!inWorld(d, q)) continue;
            const uint32_t j = cellIndex(d, q);
            const float z = FLD(d, 2u * j, 2);
            if (z > 0.0f) tension = fmaxf(tension, d.eprops[wType(d.cells[4ull * j])].tension);
            else if (z < 0.0f) contact += 2.0f * d.eprops[wType(d.cells[4ull * j])].wetting - 1.0f;
        }
        if (tension <= 0.0f) continue;
        float gx = 0.0f, gy = 0.0f, hx = 0.0f, hy = 0.0f, hz = 0.0f;
        for (int y = -2; y <= 2; y++)
            for (int x = -2; x <= 2; x++) {
                const float kx = (float)x, ky = (float)y;
                const float wt = wexpf(-0.5f * (kx * kx + ky * ky)) * outline(d, I2(c.x + x, c.y + y));
                gx += kx * wt;
                gy += ky * wt;
                hx += (kx * kx - 1.0f) * wt;
                hy += (ky * ky - 1.0f) * wt;
                hz += (kx * ky) * wt;
            }
        const float g2 = gx * gx + gy * gy;
        float kappa = 0.0f;
        if (g2 >= 1e-6f) kappa = -(hx * gy * gy - 2.0f * gx * gy * hz + hy * gx * gx) / (g2 * sqrtf(g2));
        const float curved = wclamp(kappa * CURVATURE_GAIN, -MAX_CURVATURE, MAX_CURVATURE);
        FLD(d, 2u * i, 0) = waterTension(d) * tension * (curved - CONTACT * contact);
    }
}

// setFaces (:1936-1994)
__device__ __forceinline__ float obstacleVel(const Dev &d, i2 c, int axis) {
    if (!inWorld(d, c)) return 0.0f;
    const uint32_t held = occupantAt(d, cellIndex(d, c));
    if (isOccupant(held)) {
        const v2 u = occupantVel(held);
        return axis == 0 ? u.x : u.y;
    }
    uint4 v = *(const uint4 *)&d.cells[4ull * cellIndex(d, c)];
    uint32_t k = kindOf(d, wType(v.x));
    if (k == KIND_SOLID || k == KIND_GAS) return 0.0f;
    v2 u = unpack2x16(v.z);
    return (axis == 0 ? u.x : u.y) * POWDER_PUSH;
}
__device__ __forceinline__ float faceViscosity(float4 ga, float4 gb) {
    if (ga.z < 0.0f || gb.z < 0.0f) return 0.0f;
    float fa = ga.z > 0.0f ? 1.0f : 0.0f, fb = gb.z > 0.0f ? 1.0f : 0.0f;
    return (fa * ga.w + fb * gb.w) / fmaxf(fa + fb, 1.0f);
}
// wasLiquid (:3315-3317) and lastPush (:3305-3310): last step's pressure
// push across a face
__device__ __forceinline__ bool wasLiquid(const Dev &d, i2 c) {
    return (wu32(FLD(d, 2u * cellIndex(d, c) + 1u, 3)) & stepBit(d, 1u)) != 0u;
}
__device__ __forceinline__ float lastPush(const Dev &d, i2 a, i2 at, float4 ga, float4 gb) {
    if ((ga.z > 0.0f && !wasLiquid(d, a)) || (gb.z > 0.0f && !wasLiquid(d, at))) return 0.0f;
    return fineBeta(ga.z, gb.z, densityAt(d, a), densityAt(d, at)) * (gb.x - ga.x);
}
// a face of a free body's cell: the body's velocity there going into the
// solve, the mean between two bodies' cells (:1841-1858)
__device__ void setBodyFace(const Dev &d, uint32_t f, i2 at, bool isU, float4 ga, float4 gb, uint32_t ra, uint32_t rb, bool drag) {
    const int axis = isU ? 0 : 1;
    const v2 x = faceMiddle(at, isU);
    float now, u;
    if (ra != WNONE && rb != WNONE) {
        const v2 ma = bodyMoving(d, ra, x, drag), mb = bodyMoving(d, rb, x, drag), va = rigidVel(d, ra, x), vb = rigidVel(d, rb, x);
        now = 0.5f * ((axis ? ma.y : ma.x) + (axis ? mb.y : mb.x));
        u = 0.5f * ((axis ? va.y : va.x) + (axis ? vb.y : vb.x));
    } else {
        const uint32_t b = ra != WNONE ? ra : rb;
        const v2 m = bodyMoving(d, b, x, drag), v = rigidVel(d, b, x);
        now = axis ? m.y : m.x;
        u = axis ? v.y : v.x;
    }
    const float4 other = ra != WNONE ? gb : ga;
    const float nu = (ra == WNONE || rb == WNONE) && other.z > 0.0f ? other.w : 0.0f;
    *(float4 *)&FLD(d, faceSlot(d, f), 0) = make_float4(now, u, 1.0f, nu);
    FLD(d, viscRhsSlot(d, f), f % 4u) = now;
}
__device__ void setFace(const Dev &d, i2 at, bool isU) {
    uint32_t f = faceIndex(d, at, isU);
    i2 a = isU ? I2(at.x - 1, at.y) : I2(at.x, at.y - 1);
    float4 ga = gridAt4(d, a), gb = gridAt4(d, at);
    int32_t weight = (int32_t)SUM(d, f, 1);
    if (d.body_on) {
        const uint32_t ra = rigidAt(d, a), rb = rigidAt(d, at);
        if (ra != WNONE || rb != WNONE) {
            setBodyFace(d, f, at, isU, ga, gb, ra, rb, false);
            return;
        }
    }
    float u = 0.0f, held = 0.0f;
    if (ga.z < 0.0f && gb.z < 0.0f) {
        held = -1.0f;
    } else if (ga.z < 0.0f || gb.z < 0.0f) {
        u = obstacleVel(d, ga.z < 0.0f ? a : at, isU ? 0 : 1);
        held = 1.0f;
    } else if (weight > 0) {
        u = (float)(int32_t)SUM(d, f, 0) / (float)weight;
        held = 1.0f;
    }
    float now = u;
    if (held > 0.0f && d.air_on) now += surfaceDrag(d, a, at, isU, ga, gb, u);   // the wind's drag
    // (liquids all fall alike, under the Earth's pull and the masses', along
    // both axes as it is at the face, :1968-1972)
    if (held > 0.0f && ga.z >= 0.0f && gb.z >= 0.0f) {
        const v2 pull = pullAt(d, isU ? V2((float)at.x, (float)at.y + 0.5f) : V2((float)at.x + 0.5f, (float)at.y));
        now += isU ? pull.x : pull.y;
    }
    float start = now;
    const float nu = faceViscosity(ga, gb);
    if (nu > 0.0f) {            // isViscousFace: last step's push off, the sweeps' last change on (:1975-1978)
        now -= lastPush(d, a, at, ga, gb);
        start = now + FLD(d, viscLastSlot(d, f), f % 4u);
    }
    *(float4 *)&FLD(d, faceSlot(d, f), 0) = make_float4(start, u, held, nu);
    FLD(d, viscRhsSlot(d, f), f % 4u) = now;
}
__global__ void __launch_bounds__(64) k_setFaces(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (!inWorld(d, c)) continue;
        Face3 fs[4];
        facesOf(d, c, fs);
        for (int j = 0; j < 4; j++)
            if (fs[j].isU >= 0) setFace(d, I2(fs[j].x, fs[j].y), fs[j].isU == 1);
    }
}

// viscousFace (:3361-3403): a red-black Gauss-Seidel step of implicit
// viscosity on the face at `at` if it is of `color` (no bodies: heavy 1)
#define VISC_OMEGA 1.6f
__device__ void viscousFace(const Dev &d, i2 at, bool isU, int color) {
    if (((at.x + at.y) & 1) != color) return;
    const i2 a = isU ? I2(at.x - 1, at.y) : I2(at.x, at.y - 1);
    const float4 ga = gridAt4(d, a), gb = gridAt4(d, at);
    // a body's face with liquid beyond it is solved too, the body's mass on
    // its side (:3361-3377)
    const uint32_t ra = rigidAt(d, a), rb = rigidAt(d, at);
    const bool surface = (ra != WNONE) != (rb != WNONE) && (ra != WNONE ? gb.z : ga.z) > 0.0f;
    if (!surface && (ga.z < 0.0f || gb.z < 0.0f || ra != WNONE || rb != WNONE)) return;
    const uint32_t f = faceIndex(d, at, isU);
    const uint64_t fs = faceSlot(d, f);
    const float nu = FLD(d, fs, 3);
    if (nu <= 0.0f) return;
    float heavy = 1.0f;
    if (surface) {
        const float liquid = ra != WNONE ? gb.z : ga.z;
        heavy = 2.0f * liquid / (liquid + bodyDensity(d, ra != WNONE ? ra : rb));
    }
    const float through[2] = {ga.z > 0.0f ? ga.w : 0.0f, gb.z > 0.0f ? gb.w : 0.0f};
    // the four faces beside along the same axis, loaded at once
    float4 other[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const i2 o = beside(at, k);
        const bool in = faceInWorld(d, o, isU);
        const float4 v = *(const float4 *)&FLD(d, faceSlot(d, faceIndex(d, in ? o : at, isU)), 0);
        other[k] = in ? v : make_float4(0.0f, 0.0f, -1.0f, 0.0f);
    }
    float sum = 0.0f, n = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bool across = isU == (k >= 2);
        float kk = (k & 1) ? through[1] : through[0];
        if (across) kk = other[k].z < 0.0f ? nu : 2.0f * nu * other[k].w / fmaxf(nu + other[k].w, 1e-12f);
        if (other[k].z < 0.0f) {
            n += kk * (across ? 2.0f : 1.0f);
        } else if (other[k].z > 0.0f) {
            sum += kk * other[k].x;
            n += kk;
        }
    }
    const float rhs = FLD(d, viscRhsSlot(d, f), f % 4u);
    FLD(d, fs, 0) = wmix(FLD(d, fs, 0), (rhs + heavy * sum) / (1.0f + heavy * n), VISC_OMEGA);
}
// viscosityRed / viscosityBlack (color 0, 1) and unpress (2) (:3405-3466):
// over the blocks worked on with viscous liquid in or beside them, each
// cell's faces.  unpress: on the faces the sweeps solve, what they did kept
// for the next step (viscLast) and last step's push back on; on the rest,
// viscLast cleared.
__global__ void __launch_bounds__(64) k_viscosity(WorldGrid g, int color) {
    WORLD(g);
    if (!d.full || (color < 2 && !d.visc_on)) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        const i2 b = activeBlockAt(d, blk_);
        uint32_t near = 0u;
        for (int32_t y = b.y - 1; y <= b.y + 1; y++)
            for (int32_t x = b.x - 1; x <= b.x + 1; x++)
                if (inBlocks(d, I2(x, y))) near |= d.blocks[STATE_SLOT(d, blockIndex(d, I2(x, y)))];
        if ((near & HAS_VISCOUS) == 0u) continue;
        const i2 c = blockCell(b, threadIdx.x);
        if (!inWorld(d, c)) continue;
        Face3 fs[4];
        facesOf(d, c, fs);
        for (int j = 0; j < 4; j++) {
            if (fs[j].isU < 0) continue;
            const i2 at = I2(fs[j].x, fs[j].y);
            const bool isU = fs[j].isU == 1;
            if (color < 2) {
                viscousFace(d, at, isU, color);
                continue;
            }
            const i2 a = isU ? I2(at.x - 1, at.y) : I2(at.x, at.y - 1);
            const float4 ga = gridAt4(d, a), gb = gridAt4(d, at);
            const uint32_t f = faceIndex(d, at, isU);
            if (rigidAt(d, a) == WNONE && rigidAt(d, at) == WNONE && faceViscosity(ga, gb) > 0.0f) {
                const uint64_t fs0 = faceSlot(d, f);
                FLD(d, viscLastSlot(d, f), f % 4u) = FLD(d, fs0, 0) - FLD(d, viscRhsSlot(d, f), f % 4u);
                FLD(d, fs0, 0) += lastPush(d, a, at, ga, gb);
            } else if (FLD(d, viscLastSlot(d, f), f % 4u) != 0.0f) {
                FLD(d, viscLastSlot(d, f), f % 4u) = 0.0f;
            }
        }
    }
}

// divergence (:1996-2025)
__global__ void __launch_bounds__(64) k_divergence(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    if (bid == 0 && threadIdx.x == 0) {
        d.blocks[COUNT_SLOT(d, COUNT_RESIDUAL)] = 0u;
        d.blocks[COUNT_SLOT(d, COUNT_CYCLES)] = 0u;
    }
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (!inWorld(d, c)) continue;
        uint32_t i = cellIndex(d, c);
        if (FLD(d, 2u * i, 2) <= 0.0f) continue;
        float out = FLD(d, faceSlot(d, uFace(d, c.x + 1, c.y)), 0) - FLD(d, faceSlot(d, uFace(d, c.x, c.y)), 0)
            + FLD(d, faceSlot(d, vFace(d, c.x, c.y + 1)), 0) - FLD(d, faceSlot(d, vFace(d, c.x, c.y)), 0);
        FLD(d, 2u * i, 1) = out;
        float n = densityAt(d, c);
        float room = 1.0f;                                   // posedRoom (:2035-2045)
        if (d.body_on)
            for (int k = 0; k < 4; k++) {
                const i2 q = beside(c, k);
                const uint32_t b = rigidAt(d, q);
                if (b != WNONE) room += 0.5f - bodyReach(d, b, I2(c.x - q.x, c.y - q.y));
            }
        room = wclamp(room, 0.1f, 1.9f);
        FLD(d, 2u * i + 1u, 1) = FLD(d, 2u * i + 1u, 2) < 0.0f ? 0.0f : -SPREAD * (n - room);
    }
}

// ============================================================ the solve
__device__ __forceinline__ bool inLevel(const Dev &d, uint32_t l, i2 c) {
    return c.x >= 0 && c.y >= 0 && c.x < (int32_t)d.lw[l] && c.y < (int32_t)d.lh[l];
}
__device__ __forceinline__ uint64_t levelSlot(const Dev &d, uint32_t l, i2 c) {
    return d.lstart[l] + 2ull * ((uint64_t)c.y * d.lw[l] + (uint64_t)c.x);
}
__device__ __forceinline__ float rhoAt(const Dev &d, uint32_t l, i2 c) {     // :2251-2259
    if (!inLevel(d, l, c)) return -1.0f;
    if (l == 0) return FLD(d, 2u * cellIndex(d, c), 2);
    return FLD(d, levelSlot(d, l, c) + 1u, 0);
}
__device__ __forceinline__ v2 potAt(const Dev &d, uint32_t l, i2 c) {        // :2262-2271
    if (!inLevel(d, l, c)) return V2(0.0f, 0.0f);
    if (l == 0) { uint32_t i = 2u * cellIndex(d, c); return V2(FLD(d, i, 0), FLD(d, i + 1u, 0)); }
    uint64_t at = levelSlot(d, l, c);
    return V2(FLD(d, at, 0), FLD(d, at, 1));
}
__device__ __forceinline__ v2 rhsAt(const Dev &d, uint32_t l, i2 c) {
    if (l == 0) { uint32_t i = 2u * cellIndex(d, c); return V2(FLD(d, i, 1), FLD(d, i + 1u, 1)); }
    uint64_t at = levelSlot(d, l, c);
    return V2(FLD(d, at, 2), FLD(d, at, 3));
}
// The V-cycle's storage by S (STORE_*): float in the reference's layouts
// (pa, pb, r2; the field's level records), or two bytes a value in the
// arrays beside them.  Level l >= 1's cell c holds p (px, py), the rhs
// (rx, ry) and rho; the fine level's buffers (0 pa, 1 pb, 2 r2) two values.
__device__ __forceinline__ uint64_t lvIndex(const Dev &d, uint32_t l, i2 c) { return (uint64_t)c.y * d.lw[l] + (uint64_t)c.x; }
template <int S> __device__ __forceinline__ float4 lvLoad(const Dev &d, uint32_t l, i2 c) {
    if (S == STORE_F32) return *(const float4 *)&FLD(d, levelSlot(d, l, c), 0);
    uint2 u = *(const uint2 *)(d.vlev[l] + 4ull * lvIndex(d, l, c));
    return make_float4(dec16<S>((uint16_t)u.x), dec16<S>((uint16_t)(u.x >> 16)), dec16<S>((uint16_t)u.y), dec16<S>((uint16_t)(u.y >> 16)));
}
template <int S> __device__ __forceinline__ float2 lvLoadP(const Dev &d, uint32_t l, i2 c) {
    if (S == STORE_F32) return *(const float2 *)&FLD(d, levelSlot(d, l, c), 0);
    uint32_t u = *(const uint32_t *)(d.vlev[l] + 4ull * lvIndex(d, l, c));
    return make_float2(dec16<S>((uint16_t)u), dec16<S>((uint16_t)(u >> 16)));
}
template <int S> __device__ __forceinline__ float lvRho(const Dev &d, uint32_t l, i2 c) {
    if (S == STORE_F32) return FLD(d, levelSlot(d, l, c) + 1u, 0);
    return d.vrho[l][lvIndex(d, l, c)];
}
template <int S> __device__ __forceinline__ void lvStore(const Dev &d, uint32_t l, i2 c, float4 v) {
    if (S == STORE_F32) {
        *(float4 *)&FLD(d, levelSlot(d, l, c), 0) = v;
        return;
    }
    uint2 u = make_uint2((uint32_t)enc16<S>(v.x) | ((uint32_t)enc16<S>(v.y) << 16), (uint32_t)enc16<S>(v.z) | ((uint32_t)enc16<S>(v.w) << 16));
    *(uint2 *)(d.vlev[l] + 4ull * lvIndex(d, l, c)) = u;
}
template <int S> __device__ __forceinline__ void lvStoreP(const Dev &d, uint32_t l, i2 c, float2 p) {
    if (S == STORE_F32) {
        *(float2 *)&FLD(d, levelSlot(d, l, c), 0) = p;
        return;
    }
    *(uint32_t *)(d.vlev[l] + 4ull * lvIndex(d, l, c)) = (uint32_t)enc16<S>(p.x) | ((uint32_t)enc16<S>(p.y) << 16);
}
template <int S> __device__ __forceinline__ void lvStoreRho(const Dev &d, uint32_t l, i2 c, float rho) {
    if (S == STORE_F32) *(float4 *)&FLD(d, levelSlot(d, l, c) + 1u, 0) = make_float4(rho, 0.0f, 0.0f, 0.0f);
    else d.vrho[l][lvIndex(d, l, c)] = rho;
}
template <int S> __device__ __forceinline__ float2 fnLoad(const Dev &d, int buf, uint32_t i) {
    if (S == STORE_F32) return (buf == 0 ? d.pa : buf == 1 ? d.pb : d.r2)[i];
    uint32_t u = (buf == 0 ? d.ha : buf == 1 ? d.hb : d.hr2)[i];
    return make_float2(dec16<S>((uint16_t)u), dec16<S>((uint16_t)(u >> 16)));
}
template <int S> __device__ __forceinline__ void fnStore(const Dev &d, int buf, uint32_t i, float2 v) {
    if (S == STORE_F32) (buf == 0 ? d.pa : buf == 1 ? d.pb : d.r2)[i] = v;
    else (buf == 0 ? d.ha : buf == 1 ? d.hb : d.hr2)[i] = (uint32_t)enc16<S>(v.x) | ((uint32_t)enc16<S>(v.y) << 16);
}
template <int S> __device__ __forceinline__ void fnStoreX(const Dev &d, int buf, uint32_t i, float x) {
    if (S == STORE_F32) (buf == 0 ? d.pa : buf == 1 ? d.pb : d.r2)[i].x = x;
    else ((uint16_t *)(buf == 0 ? d.ha : buf == 1 ? d.hb : d.hr2))[2ull * i] = enc16<S>(x);
}

struct Stencil { v2 sum; float diag; };
__device__ Stencil stencil(const Dev &d, uint32_t l, i2 c, float rho) {      // :2286-2302
    Stencil s;
    s.sum = V2(0.0f, 0.0f);
    s.diag = 0.0f;
    float dc = l == 0 ? densityAt(d, c) : 0.0f;
    for (int k = 0; k < 4; k++) {
        i2 n = beside(c, k);
        float r = rhoAt(d, l, n);
        if (r < 0.0f) continue;
        float beta = l == 0 ? fineBeta(rho, r, dc, densityAt(d, n)) : faceBeta(rho, r);
        v2 p = potAt(d, l, n);
        s.sum.x += beta * p.x;
        s.sum.y += beta * p.y;
        s.diag += beta;
    }
    return s;
}
__device__ v2 residualAt(const Dev &d, uint32_t l, i2 c) {                  // :2306-2315
    float rho = rhoAt(d, l, c);
    if (rho <= 0.0f) return V2(0.0f, 0.0f);
    Stencil s = stencil(d, l, c, rho);
    float h2 = (float)(1u << (2u * l));
    v2 pin = (l == 0 && pinnedAt(d, cellIndex(d, c))) ? V2(PINF, 0.0f) : V2(0.0f, 0.0f);
    v2 pot = potAt(d, l, c), rhs = rhsAt(d, l, c);
    return V2(rhs.x - (s.sum.x - (s.diag + pin.x) * pot.x) / h2, rhs.y - (s.sum.y - (s.diag + pin.y) * pot.y) / h2);
}

// relax / relaxShifted (:2080-2192): a 10 x 10 tile in LDS, its rim frozen
// as relaxRims snapshot it, sweep_pairs red-black pairs, one write back.
// The pre-relaxation reads pa and writes pb, the post-relaxation reads pb
// plus prolongFine's correction (:2500-2512) and writes pa: the buffer read
// is the rim snapshot, and every fluid cell of the tiles is written.
template <int S> __device__ v2 prolonged(const Dev &d, uint32_t l, i2 c);
// (entry: the active list's entry at wid, read ahead; clamped into the list)
__device__ __forceinline__ i2 tileBlock(const Dev &d, uint32_t na, uint32_t wid, int shift, uint32_t entry) {
    if (wid < na) return activeBlockAt(d, entry);
    int32_t wide = (int32_t)d.bw, high = (int32_t)(d.B / d.bw);
    int32_t j = (int32_t)wid - (int32_t)d.B;
    if (shift == 0 || j < 0 || j > high + wide) return I2(-1, -1);
    return j <= high ? I2(wide, j) : I2(j - high - 1, high);
}
// prolonged's weighted sums (:2458-2497) over its four coarse cells in its
// order (the cell under, the one across in x, in y, the diagonal; weights
// 9, 3, 3, 1), from their rho and p loaded beforehand: a cell of rho < 0
// is skipped, one of rho 0 adds its weight only.  Selects in place of the
// loop's branches: the same operations in the same order.
struct ProlongSum { float sx, sy, weight; };
__device__ __forceinline__ ProlongSum prolongSum(const float r[4], const float px[4], const float py[4]) {
    ProlongSum s = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float wk = k == 0 ? 9.0f : (k == 3 ? 1.0f : 3.0f);
        const bool on = !(r[k] < 0.0f), add = r[k] > 0.0f;
        s.weight = on ? s.weight + wk : s.weight;
        s.sx = add ? s.sx + wk * px[k] : s.sx;
        s.sy = add ? s.sy + wk * py[k] : s.sy;
    }
    return s;
}
// a wavefront's own LDS phase ended: its lanes' LDS writes seen by its lanes
__device__ __forceinline__ void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
#define NT RELAX_NT
template <int S> __global__ void __launch_bounds__(256) k_relax(WorldGrid g, int post) {
    WORLD(g);
    if (d.store != S) return;
    // the first tiles' list entries, read beside the iteration's words
    const uint32_t first = bid * WAVES * NT + (threadIdx.x / 64u) * NT;
    uint32_t entry[NT];
#pragma unroll
    for (int j = 0; j < NT; j++) entry[j] = d.blocks[ACTIVE_SLOT(d, min(first + (uint32_t)j, d.B - 1u))];
    const Iter it = iterWords(d);
    if (!it.encoded || it.converged) return;
    const int shift = ((d.shifted_first ^ (2u * it.cycle + (uint32_t)post)) & 1u) ? WIG_BLOCK / 2 : 0;
    // A workgroup takes 4 x NT tiles at a time, NT a wavefront, each in LDS
    // of its own: no wavefront reads another's, so each phase ends on the
    // wavefront's own sync, and the four go at their own pace.  Every load
    // of the NT tiles is made at once and unconditionally (clamped
    // addresses, values chosen after), so that they are in flight together.
    __shared__ float sPxA[WAVES * NT][RIM * RIM], sPyA[WAVES * NT][RIM * RIM];
    __shared__ float cRhoA[WAVES * NT][64], cPxA[WAVES * NT][64], cPyA[WAVES * NT][64];
    const uint32_t li = threadIdx.x % 64u, wave = threadIdx.x / 64u;
    const uint32_t na = it.active, edge = shift ? d.bw + d.bh + 1u : 0u;
    const int t = (int)(li / 8u + 1u) * RIM + (int)(li % 8u) + 1;
    const int32_t lw1 = (int32_t)d.lw[1], lh1 = (int32_t)d.lh[1];
    // the tiles: the active list's blocks, then (shifted) the edge tiles,
    // workgroup indices B .. B + edge of the reference's dispatch
    for (uint32_t base = bid * WAVES * NT; base < na + edge; base += nbid * WAVES * NT) {
        bool live[NT];
        i2 origin[NT];
        uint32_t bits[NT], mf[NT], mi[NT], rf[NT][2];
        float4 cf[NT];
        float2 r[NT], lp[NT], rv[NT][2];
        float lr[NT];
#pragma unroll
        for (int j = 0; j < NT; j++) {
            const uint32_t v = base + wave * NT + (uint32_t)j;
            const uint32_t wid = v < na ? v : d.B + (v - na);
            if (v != first + (uint32_t)j) entry[j] = d.blocks[ACTIVE_SLOT(d, min(v, d.B - 1u))];
            i2 b = v < na + edge ? tileBlock(d, na, wid, shift, entry[j]) : I2(-1, -1);
            origin[j] = I2(b.x * WIG_BLOCK - shift, b.y * WIG_BLOCK - shift);
            live[j] = !(b.x < 0 || origin[j].x >= d.W || origin[j].y >= d.H);
        }
#pragma unroll
        for (int j = 0; j < NT; j++) {
            const i2 o = origin[j];
            // the states of the (up to 2 x 2) blocks under the tile
            const i2 b0 = idiv8(o), b1 = idiv8(I2(o.x + WIG_BLOCK - 1, o.y + WIG_BLOCK - 1));
            bits[j] = 0u;
            for (int q = 0; q < 4; q++) {
                const i2 bq = I2(q & 1 ? b1.x : b0.x, q & 2 ? b1.y : b0.y);
                const bool use = live[j] && inBlocks(d, bq) && (q != 1 || b1.x != b0.x) && (q != 2 || b1.y != b0.y) && (q != 3 || (b1.x != b0.x && b1.y != b0.y));
                const uint32_t sb = d.blocks[STATE_SLOT(d, inBlocks(d, bq) ? blockIndex(d, bq) : 0u)];
                bits[j] |= use ? sb : 0u;
            }
            // this thread's cell
            const i2 me = I2(o.x + (int32_t)(li % 8u), o.y + (int32_t)(li / 8u));
            const bool meIn = live[j] && inWorld(d, me);
            mi[j] = packIndex(d, clampTo(me, d.W, d.H));
            const uint32_t f = d.cflag[mi[j]];
            const float4 c4 = d.coef[mi[j]];
            const float2 r2 = fnLoad<S>(d, 2, mi[j]);
            mf[j] = meIn ? f : 0u;
            cf[j] = c4;
            r[j] = r2;
            // level 1 from 8 x 8 cells one before the rim's first, for
            // prolongFine's correction
            const int32_t cx0 = ((o.x - 1) >> 1) - 1, cy0 = ((o.y - 1) >> 1) - 1;
            lr[j] = -1.0f;
            lp[j] = make_float2(0.0f, 0.0f);
            if (post) {
                const i2 cc = I2(cx0 + (int32_t)(li % 8u), cy0 + (int32_t)(li / 8u)), ck = clampTo(cc, lw1, lh1);
                const float rr = lvRho<S>(d, 1u, ck);
                const float2 pp = lvLoadP<S>(d, 1u, ck);
                const bool ccIn = inLevel(d, 1u, cc);
                lr[j] = ccIn ? rr : -1.0f;
                lp[j] = ccIn ? pp : make_float2(0.0f, 0.0f);
            }
            // the rim buffer's cells (two a thread)
            for (uint32_t q = 0; q < 2u; q++) {
                const uint32_t k = li + 64u * q;
                const i2 c = I2(o.x - 1 + (int32_t)(k % RIM), o.y - 1 + (int32_t)(k / RIM));
                const bool cIn = live[j] && k < RIM * RIM && inWorld(d, c);
                const uint32_t ci = packIndex(d, clampTo(c, d.W, d.H));
                const float2 vv = fnLoad<S>(d, post ? 1 : 0, ci);
                const uint32_t ff = d.cflag[ci];
                rv[j][q] = cIn ? vv : make_float2(0.0f, 0.0f);
                rf[j][q] = (cIn && post) ? ff : 0u;
            }
        }
        if (post) {
#pragma unroll
            for (int j = 0; j < NT; j++) {
                const uint32_t s_ = wave * NT + (uint32_t)j;
                cRhoA[s_][li] = lr[j];
                cPxA[s_][li] = lp[j].x;
                cPyA[s_][li] = lp[j].y;
            }
            waveSync();
        }
#pragma unroll
        for (int j = 0; j < NT; j++) {
            const uint32_t s_ = wave * NT + (uint32_t)j;
            const i2 o = origin[j];
            const int32_t cx0 = ((o.x - 1) >> 1) - 1, cy0 = ((o.y - 1) >> 1) - 1;
            for (uint32_t q = 0; q < 2u; q++) {
                uint32_t k = li + 64u * q;
                if (k >= RIM * RIM) continue;
                i2 c = I2(o.x - 1 + (int32_t)(k % RIM), o.y - 1 + (int32_t)(k / RIM));
                float2 v2_ = rv[j][q];
                if (post && (rf[j][q] & CF_FLUID)) {
                    // prolonged() (:2458-2497) from the staged cells
                    i2 coarse = I2(c.x / 2, c.y / 2);
                    i2 sg = I2((c.x & 1) == 1 ? 1 : -1, (c.y & 1) == 1 ? 1 : -1);
                    float rc[4], cx[4], cy[4];
#pragma unroll
                    for (int o4 = 0; o4 < 4; o4++) {
                        i2 oo = o4 == 0 ? I2(0, 0) : (o4 == 1 ? I2(sg.x, 0) : (o4 == 2 ? I2(0, sg.y) : sg));
                        int32_t jj = (coarse.y + oo.y - cy0) * 8 + (coarse.x + oo.x - cx0);
                        rc[o4] = cRhoA[s_][jj];
                        cx[o4] = cPxA[s_][jj];
                        cy[o4] = cPyA[s_][jj];
                    }
                    const ProlongSum ps = prolongSum(rc, cx, cy);
                    float dn = fmaxf(ps.weight, 1e-6f);
                    v2_.x = v2_.x + ps.sx / dn;
                    v2_.y = v2_.y + ps.sy / dn;
                }
                sPxA[s_][k] = v2_.x;
                sPyA[s_][k] = v2_.y;
            }
        }
        waveSync();
        float b0[NT], b1[NT], b2[NT], b3[NT], rhsx[NT], rhsy[NT], diag[NT], dx[NT], dy[NT];
        bool fluid[NT], solving[NT];
        int color[NT];
#pragma unroll
        for (int j = 0; j < NT; j++) {
            live[j] = live[j] && (bits[j] & HAS_FLUID) != 0u;
            fluid[j] = live[j] && (mf[j] & CF_FLUID) != 0u;
            const bool pinned = fluid[j] && (mf[j] & CF_PINNED) != 0u;
            b0[j] = fluid[j] ? fabsf(cf[j].x) : 0.0f;
            b1[j] = fluid[j] ? fabsf(cf[j].y) : 0.0f;
            b2[j] = fluid[j] ? fabsf(cf[j].z) : 0.0f;
            b3[j] = fluid[j] ? fabsf(cf[j].w) : 0.0f;
            rhsx[j] = fluid[j] ? r[j].x : 0.0f;
            rhsy[j] = fluid[j] ? r[j].y : 0.0f;
            diag[j] = b0[j] + b1[j] + b2[j] + b3[j];
            dx[j] = diag[j] + (pinned ? PINF : 0.0f);
            dy[j] = fmaxf(diag[j], 1e-30f);
            solving[j] = dx[j] > 0.0f;
            color[j] = (origin[j].x + (int32_t)(li % 8u) + origin[j].y + (int32_t)(li / 8u)) & 1;
        }
        for (int sw = 0; sw < 2 * (int)d.sweep_pairs; sw++) {
#pragma unroll
            for (int j = 0; j < NT; j++) {
                if (solving[j] && color[j] == (sw & 1)) {
                    float *px = sPxA[wave * NT + (uint32_t)j], *py = sPyA[wave * NT + (uint32_t)j];
                    float sx = b0[j] * px[t - 1] + b1[j] * px[t + 1] + b2[j] * px[t - RIM] + b3[j] * px[t + RIM];
                    float sy = b0[j] * py[t - 1] + b1[j] * py[t + 1] + b2[j] * py[t - RIM] + b3[j] * py[t + RIM];
                    float nx = wmix(px[t], (sx - rhsx[j]) / dx[j], SMOOTH_OMEGA);
                    float ny = wmix(py[t], (sy - rhsy[j]) / dy[j], SMOOTH_OMEGA);
                    px[t] = nx;
                    if (diag[j] > 0.0f) py[t] = ny;
                }
            }
            waveSync();
        }
#pragma unroll
        for (int j = 0; j < NT; j++)
            if (fluid[j]) fnStore<S>(d, post ? 0 : 1, mi[j], make_float2(sPxA[wave * NT + (uint32_t)j][t], sPyA[wave * NT + (uint32_t)j][t]));
        waveSync();
    }
}

// restrictLevel into level 1 (:2342-2377), each iteration: rhs from the
// fluid children's residuals, a quarter of their sum; rho as k_levelRho put
// it at the solve's start.  The fine residual (residualAt, :2306-2315) from
// the packed arrays: p as the pre-relaxation left it (pb), the betas (0
// where the stencil skips the neighbour), the pin, the rhs.  A workgroup
// takes 8 x 8 cells of level 1 over 16 x 16 fine cells, their pb and its
// ring staged in LDS, every load made at once.
#define RF 18
template <int S> __global__ void __launch_bounds__(64) k_restrict(WorldGrid g) {
    WORLD(g);
    if (d.store != S) return;
    const Iter it = iterWords(d);
    const uint32_t gx_ = (d.lw[1] + 7u) / 8u, bx = bid % gx_, by = bid / gx_;
    __shared__ float sPx[RF * RF], sPy[RF * RF];
    const uint32_t li = threadIdx.x;
    const int32_t fx0 = (int32_t)bx * 16 - 1, fy0 = (int32_t)by * 16 - 1;
    // level 1's rho, and whether any of the four fine blocks under the tile
    // holds fluid: level 1's rho is <= 0 over the others (0 over blocks not
    // worked on, k_levelRho; from fine rho <= 0 elsewhere), and a cell of
    // rho <= 0 restricts 0
    const i2 c = I2((int32_t)(bx * 8u + li % 8u), (int32_t)(by * 8u + li / 8u));
    const bool in = inLevel(d, 1u, c);
    // The loads at addresses always there, their values chosen after: the
    // tile's rho and block states, then (a tile with fluid) its fine cells.
    // One world makes the first beside the iteration's words (in a batch,
    // worlds that have converged would make them for nothing).
    float rho = 0.0f;
    uint32_t st = 0;
    float2 pv[6];
    uint32_t kf[4];
    float4 kc[4];
    float2 kr[4];
    auto loadTile = [&]() {
        const float rr = lvRho<S>(d, 1u, clampTo(c, (int32_t)d.lw[1], (int32_t)d.lh[1]));
        for (int q = 0; q < 4; q++) {
            i2 fb = I2(2 * (int32_t)bx + (q & 1), 2 * (int32_t)by + (q >> 1));
            const uint32_t sb = d.blocks[STATE_SLOT(d, inBlocks(d, fb) ? blockIndex(d, fb) : 0u)];
            st |= inBlocks(d, fb) ? sb : 0u;
        }
        rho = in ? rr : 0.0f;
    };
    auto loadCells = [&]() {
        for (uint32_t q = 0; q < 6u; q++) {
            uint32_t k = li + 64u * q;
            i2 fc = I2(fx0 + (int32_t)(k % RF), fy0 + (int32_t)(k / RF));
            const float2 v = fnLoad<S>(d, 1, packIndex(d, clampTo(fc, d.W, d.H)));
            pv[q] = (k < RF * RF && inWorld(d, fc)) ? v : make_float2(0.0f, 0.0f);
        }
        for (int q = 0; q < 4; q++) {
            i2 k = I2(2 * c.x + (q & 1), 2 * c.y + (q >> 1));
            const uint32_t ki = packIndex(d, clampTo(k, d.W, d.H));
            const uint32_t f = d.cflag[ki];
            const float4 cc = d.coef[ki];
            const float2 r = fnLoad<S>(d, 2, ki);
            const bool use = rho > 0.0f && inWorld(d, k);
            kf[q] = use ? f : 0u;
            kc[q] = use ? cc : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            kr[q] = use ? r : make_float2(0.0f, 0.0f);
        }
    };
    const bool early = g.solo != 0u;
    if (early) loadTile();
    if (!it.encoded || it.converged || it.active == 0u) return;
    if (!early) loadTile();
    if ((st & HAS_FLUID) == 0u) {
        if (in) lvStore<S>(d, 1u, c, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        return;
    }
    loadCells();
    for (uint32_t q = 0; q < 6u; q++) {
        uint32_t k = li + 64u * q;
        if (k < RF * RF) {
            sPx[k] = pv[q].x;
            sPy[k] = pv[q].y;
        }
    }
    __syncthreads();
    if (!in) return;
    // the fluid children's residuals
    float sx = 0.0f, sy = 0.0f;
    if (rho > 0.0f)
        for (int q = 0; q < 4; q++) {
            if (!(kf[q] & CF_FLUID)) continue;
            const int32_t t = (2 * c.y + (q >> 1) - fy0) * RF + (2 * c.x + (q & 1) - fx0);
            const int32_t at[4] = {t - 1, t + 1, t - RF, t + RF};
            const float bk[4] = {fabsf(kc[q].x), fabsf(kc[q].y), fabsf(kc[q].z), fabsf(kc[q].w)};
            float nx = 0.0f, ny = 0.0f, diag = 0.0f;
            for (int k4 = 0; k4 < 4; k4++) {
                if (bk[k4] == 0.0f) continue;
                nx += bk[k4] * sPx[at[k4]];
                ny += bk[k4] * sPy[at[k4]];
                diag += bk[k4];
            }
            float pin = (kf[q] & CF_PINNED) ? PINF : 0.0f;
            sx += kr[q].x - (nx - (diag + pin) * sPx[t]);
            sy += kr[q].y - (ny - (diag + 0.0f) * sPy[t]);
        }
    lvStore<S>(d, 1u, c, make_float4(0.0f, 0.0f, sx * 0.25f, sy * 0.25f));
}
// level 1's rho (restrictLevel's, :2342-2377), once a solve: it depends on
// the fine rho alone
template <int S> __global__ void __launch_bounds__(64) k_levelRho(WorldGrid g) {
    WORLD(g);
    if (d.store != S) return;
    if (!d.full || d.fluid_cycles == 0u) return;
    const uint32_t gx_ = (d.lw[1] + 7u) / 8u, bx = bid % gx_, by = bid / gx_;
    if (activeCount(d) == 0u) return;
    i2 c = I2((int32_t)(bx * 8u + threadIdx.x % 8u), (int32_t)(by * 8u + threadIdx.x / 8u));
    if (!inLevel(d, 1u, c)) return;
    float rho = 0.0f;
    i2 b = I2(c.x * 2 / WIG_BLOCK, c.y * 2 / WIG_BLOCK);
    if (!(inBlocks(d, b) && (d.blocks[STATE_SLOT(d, blockIndex(d, b))] & ACTIVE) == 0u)) {
        float fluid = 0.0f, sum = 0.0f;
        bool air = false;
        for (int q = 0; q < 4; q++) {
            float rk = rhoAt(d, 0u, I2(2 * c.x + (q & 1), 2 * c.y + (q >> 1)));
            if (rk == 0.0f) air = true;
            else if (rk > 0.0f) {
                fluid += 1.0f;
                sum += rk;
            }
        }
        rho = air ? 0.0f : (fluid > 0.0f ? sum / fluid : -1.0f);
    }
    lvStoreRho<S>(d, 1u, c, rho);
}

// prolonged (:2458-2497); prolong is in k_levelUp, prolongFine in k_relax
template <int S> __device__ v2 prolonged(const Dev &d, uint32_t l, i2 c) {
    i2 coarse = I2(c.x / 2, c.y / 2);
    i2 s = I2((c.x & 1) == 1 ? 1 : -1, (c.y & 1) == 1 ? 1 : -1);
    v2 sum = V2(0.0f, 0.0f);
    float weight = 0.0f;
    for (int k = 0; k < 4; k++) {
        i2 off = k == 0 ? I2(0, 0) : (k == 1 ? I2(s.x, 0) : (k == 2 ? I2(0, s.y) : s));
        float wk = k == 0 ? 9.0f : (k == 3 ? 1.0f : 3.0f);
        i2 at = I2(coarse.x + off.x, coarse.y + off.y);
        float r = inLevel(d, l + 1u, at) ? lvRho<S>(d, l + 1u, at) : -1.0f;
        if (r < 0.0f) continue;
        weight += wk;
        if (r > 0.0f) {
            float2 p = lvLoadP<S>(d, l + 1u, at);
            sum.x += wk * p.x;
            sum.y += wk * p.y;
        }
    }
    float dn = fmaxf(weight, 1e-6f);
    return V2(sum.x / dn, sum.y / dn);
}
// A level's red-black sweeps (smoothRed, smoothBlack: :2317-2387) over one
// workgroup's T x T tile in LDS.  The tile carries a halo as wide as the
// number of half-sweeps, so each half-sweep's values are right one cell
// further in and the tile's own cells end exactly as the passes over the
// whole level leave them.  The faces' betas (faceBeta, :369-374) are taken
// once a tile: level rho does not change within a sweep.
//
// k_levelDown(l): level l as restrictLevel left it; the sweeps; then
// restrictLevel into level l + 1 from the tile's cells (one more halo cell,
// so the residuals at the tile's edge see right neighbours).
// k_levelUp(l): prolong's correction (:2459-2497) from level l + 1, staged
// in LDS; the sweeps.  Both write the tile's p back to the field.
#define TILE_THREADS 256u
static_assert(16u * 16u <= TILE_THREADS, "k_levelDown: a thread restricts at most one coarse cell (tile sides of 32 or less)");
// A thread's cells of one colour, at most SWEEP_K of them, held in registers
// across a level's half-sweeps with what no sweep changes: the four
// neighbours' betas in beside()'s order (0 where the stencil skips one),
// their sum (the diagonal) and h^2 times the rhs.  Each half-sweep then
// issues all its LDS loads at once and sums without branches.  A skipped
// neighbour's term is selected away, so the sums take the same operations
// in the same order as smoothAt's loop and the values are the same, bit for
// bit; its load reads the cell itself, an address always in the level.
#define SWEEP_K 2
struct SweepSet {
    int32_t j[SWEEP_K];   // the cell (-1: none)
    float b[SWEEP_K][4], diag[SWEEP_K], hx[SWEEP_K], hy[SWEEP_K];
    bool any[SWEEP_K];    // whether any lane of the wavefront has a cell m (the same in every lane)
};
// Each slot's wavefront vote, once the sets are made: a slot no lane of the
// wavefront uses is passed over whole, its LDS loads included (on the small
// levels most wavefronts have no cell, and their loads had taken the LDS's
// cycles all the same).
__device__ __forceinline__ void sweepVote(SweepSet &s) {
#pragma unroll
    for (int m = 0; m < SWEEP_K; m++) s.any[m] = __any(s.j[m] >= 0) != 0;
}
__device__ __forceinline__ void sweepNone(SweepSet &s, int m) {
    s.j[m] = -1;
    s.b[m][0] = s.b[m][1] = s.b[m][2] = s.b[m][3] = 0.0f;
    s.diag[m] = 1.0f;
    s.hx[m] = s.hy[m] = 0.0f;
}
// cell j with betas b0..b3 (rho > 0 there): kept if the diagonal is positive
__device__ __forceinline__ void sweepCell(SweepSet &s, int m, int32_t j, float b0, float b1, float b2, float b3, float h2, float rx, float ry) {
    float diag = 0.0f;
    diag = b0 != 0.0f ? diag + b0 : diag;
    diag = b1 != 0.0f ? diag + b1 : diag;
    diag = b2 != 0.0f ? diag + b2 : diag;
    diag = b3 != 0.0f ? diag + b3 : diag;
    if (diag <= 0.0f) return;
    s.j[m] = j;
    s.b[m][0] = b0;
    s.b[m][1] = b1;
    s.b[m][2] = b2;
    s.b[m][3] = b3;
    s.diag[m] = diag;
    s.hx[m] = h2 * rx;
    s.hy[m] = h2 * ry;
}
// one half-sweep of a set over the w-wide arrays px, py
__device__ __forceinline__ void sweepHalf(const SweepSet &s, float *px, float *py, int32_t w, float omega) {
    float ox[SWEEP_K], oy[SWEEP_K], nx[SWEEP_K][4], ny[SWEEP_K][4];
#pragma unroll
    for (int m = 0; m < SWEEP_K; m++) {
        if (!s.any[m]) continue;
        const int32_t j = s.j[m] < 0 ? 0 : s.j[m];
        const int32_t nb[4] = {j - 1, j + 1, j - w, j + w};
        ox[m] = px[j];
        oy[m] = py[j];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int32_t q = s.b[m][k] != 0.0f ? nb[k] : j;
            nx[m][k] = px[q];
            ny[m][k] = py[q];
        }
    }
#pragma unroll
    for (int m = 0; m < SWEEP_K; m++) {
        if (!s.any[m] || s.j[m] < 0) continue;
        float sx = 0.0f, sy = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float b = s.b[m][k];
            sx = b != 0.0f ? sx + b * nx[m][k] : sx;
            sy = b != 0.0f ? sy + b * ny[m][k] : sy;
        }
        px[s.j[m]] = wmix(ox[m], (sx - s.hx[m]) / s.diag[m], omega);
        py[s.j[m]] = wmix(oy[m], (sy - s.hy[m]) / s.diag[m], omega);
    }
}
struct LevelTile {
    float *px, *py, *rx, *ry, *rho, *bR, *bD;
    int32_t R, ox, oy;   // side with the halo; the halo's corner in level cells
};
__device__ __forceinline__ LevelTile levelTile(float *sm, uint32_t T, uint32_t H, uint32_t bx, uint32_t by) {
    LevelTile t;
    t.R = (int32_t)(T + 2u * H);
    uint32_t n = (uint32_t)(t.R * t.R);
    t.px = sm;
    t.py = sm + n;
    t.rx = sm + 2u * n;
    t.ry = sm + 3u * n;
    t.rho = sm + 4u * n;
    t.bR = sm + 5u * n;
    t.bD = sm + 6u * n;
    t.ox = (int32_t)(bx * T) - (int32_t)H;
    t.oy = (int32_t)(by * T) - (int32_t)H;
    return t;
}
// the tile's cells, 16 x 16 threads stepping over its rows and columns
#define FOR_TILE(t, x, y) \
    for (int32_t y = (int32_t)(threadIdx.x / 16u); y < (t).R; y += 16) \
        for (int32_t x = (int32_t)(threadIdx.x % 16u); x < (t).R; x += 16)
// The tile's cells of the kernel's level (its Dev's level 1) into LDS,
// every load made at once: TILE_LOADS a thread a round, one round for the
// tiles the LDS holds (a side with its halo of 48 or less); outside the
// level p and rhs 0, rho -1.  The correction p: with UP the down pass's
// (pdown, which no tile of this launch writes); else 0, as restrictLevel
// left every cell of the level (:2342-2377).
template <int S> __device__ __forceinline__ float2 pdownLoad(const Dev &d, i2 c) {
    const uint2 u = d.pdown[1][lvIndex(d, 1u, c)];
    if (S == STORE_F32) return make_float2(__uint_as_float(u.x), __uint_as_float(u.y));
    return make_float2(dec16<S>((uint16_t)u.x), dec16<S>((uint16_t)(u.x >> 16)));
}
template <int S> __device__ __forceinline__ void pdownStore(const Dev &d, i2 c, float2 p) {
    uint2 &u = d.pdown[1][lvIndex(d, 1u, c)];
    if (S == STORE_F32) u = make_uint2(__float_as_uint(p.x), __float_as_uint(p.y));
    else u.x = (uint32_t)enc16<S>(p.x) | ((uint32_t)enc16<S>(p.y) << 16);
}
#define TILE_LOADS 9
// BETAS: the cells' face betas too, as the solve's first iteration left them
// in the level's beta array (0 outside the level, as tileBetas leaves them)
template <int S, bool UP, bool BETAS> __device__ __forceinline__ void loadLevelTile(const Dev &d, const LevelTile &t) {
    const int32_t n = t.R * t.R, w = (int32_t)d.lw[1], h = (int32_t)d.lh[1], tid = (int32_t)threadIdx.x;
    for (int32_t b0 = 0; b0 < n; b0 += TILE_LOADS * (int32_t)TILE_THREADS) {
        float4 e[TILE_LOADS];
        float2 p[TILE_LOADS], bt[TILE_LOADS];
        float r[TILE_LOADS];
#pragma unroll
        for (int q = 0; q < TILE_LOADS; q++) {
            if (b0 + q * (int32_t)TILE_THREADS >= n) break;
            const int32_t k = min(b0 + q * (int32_t)TILE_THREADS + tid, n - 1);
            const i2 c = clampTo(I2(t.ox + k % t.R, t.oy + k / t.R), w, h);
            e[q] = lvLoad<S>(d, 1u, c);
            p[q] = UP ? pdownLoad<S>(d, c) : make_float2(0.0f, 0.0f);
            r[q] = lvRho<S>(d, 1u, c);
            bt[q] = BETAS ? d.vbeta[1][lvIndex(d, 1u, c)] : make_float2(0.0f, 0.0f);
        }
#pragma unroll
        for (int q = 0; q < TILE_LOADS; q++) {
            if (b0 + q * (int32_t)TILE_THREADS >= n) break;
            const int32_t k = b0 + q * (int32_t)TILE_THREADS + tid;
            if (k >= n) continue;
            const bool in = inLevel(d, 1u, I2(t.ox + k % t.R, t.oy + k / t.R));
            t.px[k] = in ? p[q].x : 0.0f;
            t.py[k] = in ? p[q].y : 0.0f;
            t.rx[k] = in ? e[q].z : 0.0f;
            t.ry[k] = in ? e[q].w : 0.0f;
            t.rho[k] = in ? r[q] : -1.0f;
            if (BETAS) {
                t.bR[k] = in ? bt[q].x : 0.0f;
                t.bD[k] = in ? bt[q].y : 0.0f;
            }
        }
    }
}
// the betas of each cell's right and lower faces, 0 where the stencil skips
// the neighbour (faceBeta is 0 exactly when rho > 0 and the other is < 0)
__device__ __forceinline__ void tileBetas(const LevelTile &t) {
    FOR_TILE(t, x, y) {
        int32_t j = y * t.R + x;
        float r = t.rho[j];
        t.bR[j] = x + 1 < t.R ? faceBeta(r, t.rho[j + 1]) : 0.0f;
        t.bD[j] = y + 1 < t.R ? faceBeta(r, t.rho[j + t.R]) : 0.0f;
    }
}
// The tile's own cells' betas into the level's beta array, for the solve's
// later iterations and k_levelUp.  (The rim's last column and row, whose
// betas tileBetas leaves 0, are halo: no sweep, residual or restriction reads
// a beta there, so a tile loading the array's true values there computes
// the same.)
__device__ __forceinline__ void storeTileBetas(const Dev &d, const LevelTile &t, uint32_t T, uint32_t H) {
    for (uint32_t k = threadIdx.x; k < T * T; k += TILE_THREADS) {
        const int32_t x = (int32_t)(k % T + H), y = (int32_t)(k / T + H), j = y * t.R + x;
        const i2 c = I2(t.ox + x, t.oy + y);
        if (inLevel(d, 1u, c)) d.vbeta[1][lvIndex(d, 1u, c)] = make_float2(t.bR[j], t.bD[j]);
    }
}
// smoothAt's stencil sums at tile cell j, not on the tile's rim
// (neighbours in beside()'s order; every load made at once, a skipped
// neighbour's term selected away)
struct TileSum { float sx, sy, diag; };
__device__ __forceinline__ TileSum tileStencil(const LevelTile &t, int32_t j) {
    const float b[4] = {t.bR[j - 1], t.bR[j], t.bD[j - t.R], t.bD[j]};
    const int32_t nb[4] = {j - 1, j + 1, j - t.R, j + t.R};
    float px[4], py[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        px[k] = t.px[nb[k]];
        py[k] = t.py[nb[k]];
    }
    TileSum s = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bool use = b[k] != 0.0f;
        s.sx = use ? s.sx + b[k] * px[k] : s.sx;
        s.sy = use ? s.sy + b[k] * py[k] : s.sy;
        s.diag = use ? s.diag + b[k] : s.diag;
    }
    return s;
}
__device__ void tileSweeps(const LevelTile &t, uint32_t l, uint32_t pairs) {
    float h2 = (float)(1u << (2u * l));
    const int32_t half = (t.R - 1) / 2;   // cells of one colour in a row's inner R - 2, at most
    if (t.R - 2 <= 16 * SWEEP_K) {
        // a thread's cells: rows 1 + tid / 16 + 16 m, the tid % 16-th of
        // the row's cells of the colour
        // (two sets by name, not an array: an array indexed at run time
        // would sit in scratch memory)
        SweepSet red, black;
        const int32_t col = (int32_t)(threadIdx.x % 16u);
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int m = 0; m < SWEEP_K; m++) {
                SweepSet &set = c == 0 ? red : black;
                sweepNone(set, m);
                const int32_t y = 1 + (int32_t)(threadIdx.x / 16u) + 16 * m;
                const int32_t x = 1 + ((c + t.ox + t.oy + y + 1) & 1) + 2 * col;
                if (y >= t.R - 1 || col >= half || x >= t.R - 1) continue;
                const int32_t j = y * t.R + x;
                if (t.rho[j] <= 0.0f) continue;
                sweepCell(set, m, j, t.bR[j - 1], t.bR[j], t.bD[j - t.R], t.bD[j], h2, t.rx[j], t.ry[j]);
            }
        sweepVote(red);
        sweepVote(black);
        for (uint32_t s = 0; s < pairs; s++) {
            sweepHalf(red, t.px, t.py, t.R, 1.0f);
            __syncthreads();
            sweepHalf(black, t.px, t.py, t.R, 1.0f);
            __syncthreads();
        }
        return;
    }
    for (uint32_t s = 0; s < 2u * pairs; s++) {
        int color = (int)(s & 1u);
        for (int32_t y = 1 + (int32_t)(threadIdx.x / 16u); y < t.R - 1; y += 16) {
            int32_t par = (color + t.ox + t.oy + y + 1) & 1;   // x = 1 + par + 2 i has (ox + x + oy + y) & 1 == color
            for (int32_t i = (int32_t)(threadIdx.x % 16u); i < half; i += 16) {
                int32_t x = 1 + par + 2 * i;
                if (x >= t.R - 1) continue;
                int32_t j = y * t.R + x;
                float rho = t.rho[j];
                if (rho <= 0.0f) continue;
                TileSum st = tileStencil(t, j);
                if (st.diag <= 0.0f) continue;
                float nx = (st.sx - h2 * t.rx[j]) / st.diag, ny = (st.sy - h2 * t.ry[j]) / st.diag;
                t.px[j] = wmix(t.px[j], nx, 1.0f);
                t.py[j] = wmix(t.py[j], ny, 1.0f);
            }
        }
        __syncthreads();
    }
}
template <int S> __global__ void __launch_bounds__(TILE_THREADS) k_levelDown(WorldGrid g, uint32_t l) {
    WORLD_LEVEL(g, l);
    if (d.store != S) return;
    const Iter it = iterWords(d);
    const uint32_t T = tile_;
    const uint32_t gx_ = (d.lw[1] + T - 1u) / T, bx = bid % gx_, by = bid / gx_;
    const uint8_t fluidTile = d.tflag[1][bid];
    extern __shared__ float sm[];
    const uint32_t H = 2u * d.level_pairs + 1u;
    LevelTile t = levelTile(sm, T, H, bx, by);
    float *resx = sm + 7u * (uint32_t)(t.R * t.R), *resy = resx + T * T;
    // The block states restrictLevel's shortcut reads (below), loaded
    // before the tile: a thread's coarse cell is at most one (T <= 32).
    const uint32_t T2 = T / 2u, lc = l + 1u;
    const int32_t span = 1 << lc;
    bool idleBlock = false;
    if (span <= WIG_BLOCK && threadIdx.x < T2 * T2) {
        const i2 c = I2((int32_t)(bx * T2 + threadIdx.x % T2), (int32_t)(by * T2 + threadIdx.x / T2));
        const i2 b = I2(c.x * span / WIG_BLOCK, c.y * span / WIG_BLOCK);
        const uint32_t sb = d.blocks[STATE_SLOT(d, inBlocks(d, b) ? blockIndex(d, b) : 0u)];
        idleBlock = inBlocks(d, b) && (sb & ACTIVE) == 0u;
    }
    if (!it.encoded || it.converged || it.active == 0u) return;
    // a tile none of whose own cells holds fluid leaves every value as the
    // solve's first iteration left it (only cells of rho > 0 change)
    if (it.cycle > 0u && fluidTile == 0u) return;
    // The betas: taken in the solve's first iteration, loaded after it.
    const bool first = it.cycle == 0u;
    if (first) loadLevelTile<S, false, false>(d, t);
    else loadLevelTile<S, false, true>(d, t);
    __syncthreads();
    if (first) {
        tileBetas(t);
        __syncthreads();
        storeTileBetas(d, t, T, H);
    }
    tileSweeps(t, l, d.level_pairs);
    // p back, and each tile cell's residual (residualAt, :2306-2315)
    const float ih2 = 1.0f / (float)(1u << (2u * l));   // (a power of two: x / h2 and x * ih2 are the same value)
    int anyFluid = 0;
    for (uint32_t k = threadIdx.x; k < T * T; k += TILE_THREADS) {
        int32_t x = (int32_t)(k % T + H), y = (int32_t)(k / T + H), j = y * t.R + x;
        i2 c = I2(t.ox + x, t.oy + y);
        float qx = 0.0f, qy = 0.0f;
        if (t.rho[j] > 0.0f) {
            TileSum st = tileStencil(t, j);
            qx = t.rx[j] - (st.sx - (st.diag + 0.0f) * t.px[j]) * ih2;
            qy = t.ry[j] - (st.sy - (st.diag + 0.0f) * t.py[j]) * ih2;
        }
        resx[k] = qx;
        resy[k] = qy;
        if (inLevel(d, 1u, c)) pdownStore<S>(d, c, make_float2(t.px[j], t.py[j]));
        anyFluid |= t.rho[j] > 0.0f ? 1 : 0;
    }
    anyFluid = __syncthreads_or(anyFluid);
    if (first && threadIdx.x == 0) d.tflag[1][bid] = anyFluid ? 1u : 0u;
    // restrictLevel (:2342-2377) into the tile's parents (their rho, the same
    // each iteration of a solve, stored in its first)
    for (uint32_t k = threadIdx.x; k < T2 * T2; k += TILE_THREADS) {
        i2 c = I2((int32_t)(bx * T2 + k % T2), (int32_t)(by * T2 + k / T2));
        if (!inLevel(d, 2u, c)) continue;
        float rho = 0.0f, ox = 0.0f, oy = 0.0f;
        const bool shortcut = span <= WIG_BLOCK && idleBlock;   // (k is the thread's)
        if (!shortcut) {
            float fluid = 0.0f, sum = 0.0f;
            bool air = false;
            float sx = 0.0f, sy = 0.0f;
            for (int q = 0; q < 4; q++) {
                uint32_t cx = 2u * (k % T2) + (uint32_t)(q & 1), cy = 2u * (k / T2) + (uint32_t)(q >> 1);
                float rk = t.rho[(int32_t)(cy + H) * t.R + (int32_t)(cx + H)];
                if (rk == 0.0f) air = true;
                else if (rk > 0.0f) {
                    fluid += 1.0f;
                    sum += rk;
                    sx += resx[cy * T + cx];
                    sy += resy[cy * T + cx];
                }
            }
            rho = air ? 0.0f : (fluid > 0.0f ? sum / fluid : -1.0f);
            ox = rho > 0.0f ? sx * 0.25f : 0.0f;
            oy = rho > 0.0f ? sy * 0.25f : 0.0f;
        }
        lvStore<S>(d, 2u, c, make_float4(0.0f, 0.0f, ox, oy));
        if (first) lvStoreRho<S>(d, 2u, c, rho);
    }
}
template <int S> __global__ void __launch_bounds__(TILE_THREADS) k_levelUp(WorldGrid g, uint32_t l) {
    WORLD_LEVEL(g, l);
    if (d.store != S) return;
    const Iter it = iterWords(d);
    const uint32_t T = tile_;
    const uint32_t gx_ = (d.lw[1] + T - 1u) / T, bx = bid % gx_, by = bid / gx_;
    const uint8_t fluidTile = d.tflag[1][bid];   // (k_levelDown's of this iteration's level l)
    extern __shared__ float sm[];
    const uint32_t H = 2u * d.level_pairs;
    LevelTile t = levelTile(sm, T, H, bx, by);
    // level l + 1 under the tile, from one cell before its first to one after its last
    const int32_t cx0 = (t.ox >> 1) - 1, cy0 = (t.oy >> 1) - 1, Rc = t.R / 2 + 3;
    float *cpx = sm + 7u * (uint32_t)(t.R * t.R), *cpy = cpx + Rc * Rc, *crho = cpy + Rc * Rc;
    const int32_t nc = Rc * Rc, tid = (int32_t)threadIdx.x;
    auto load = [&]() {
        // (the coarse cells' loads issued with the tile's: at most three a thread)
        float2 ce[3];
        float cr[3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int32_t k = min(tid + q * (int32_t)TILE_THREADS, nc - 1);
            const i2 c = clampTo(I2(cx0 + k % Rc, cy0 + k / Rc), (int32_t)d.lw[2], (int32_t)d.lh[2]);
            ce[q] = lvLoadP<S>(d, 2u, c);
            cr[q] = lvRho<S>(d, 2u, c);
        }
        loadLevelTile<S, true, true>(d, t);   // (the betas k_levelDown left)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int32_t k = tid + q * (int32_t)TILE_THREADS;
            if (k >= nc) continue;
            const bool in = inLevel(d, 2u, I2(cx0 + k % Rc, cy0 + k / Rc));
            cpx[k] = in ? ce[q].x : 0.0f;
            cpy[k] = in ? ce[q].y : 0.0f;
            crho[k] = in ? cr[q] : -1.0f;
        }
        for (int32_t k = tid + 3 * (int32_t)TILE_THREADS; k < nc; k += (int32_t)TILE_THREADS) {   // (sides past 48)
            const i2 c = I2(cx0 + k % Rc, cy0 + k / Rc), cc = clampTo(c, (int32_t)d.lw[2], (int32_t)d.lh[2]);
            const bool in = inLevel(d, 2u, c);
            const float2 e = lvLoadP<S>(d, 2u, cc);
            const float rr = lvRho<S>(d, 2u, cc);
            cpx[k] = in ? e.x : 0.0f;
            cpy[k] = in ? e.y : 0.0f;
            crho[k] = in ? rr : -1.0f;
        }
    };
    if (!it.encoded || it.converged || it.active == 0u) return;
    if (fluidTile == 0u) return;   // prolong and the sweeps change only cells of rho > 0
    load();
    __syncthreads();
    // prolong (:2459-2497): prolonged() from the staged coarse cells
    FOR_TILE(t, x, y) {
        int32_t j = y * t.R + x;
        i2 c = I2(t.ox + x, t.oy + y);
        float r0 = t.rho[j];
        if (!(r0 > 0.0f)) continue;     // (outside the level rho is -1)
        i2 coarse = I2(c.x / 2, c.y / 2);
        i2 sg = I2((c.x & 1) == 1 ? 1 : -1, (c.y & 1) == 1 ? 1 : -1);
        float rc[4], cx[4], cy[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            i2 o = k == 0 ? I2(0, 0) : (k == 1 ? I2(sg.x, 0) : (k == 2 ? I2(0, sg.y) : sg));
            int32_t q = (coarse.y + o.y - cy0) * Rc + (coarse.x + o.x - cx0);
            rc[k] = crho[q];
            cx[k] = cpx[q];
            cy[k] = cpy[q];
        }
        const ProlongSum ps = prolongSum(rc, cx, cy);
        float dn = fmaxf(ps.weight, 1e-6f);
        t.px[j] = t.px[j] + ps.sx / dn;
        t.py[j] = t.py[j] + ps.sy / dn;
    }
    __syncthreads();
    tileSweeps(t, l, d.level_pairs);
    for (uint32_t k = threadIdx.x; k < T * T; k += TILE_THREADS) {
        int32_t x = (int32_t)(k % T + H), y = (int32_t)(k / T + H), j = y * t.R + x;
        i2 c = I2(t.ox + x, t.oy + y);
        if (!inLevel(d, 1u, c)) continue;
        lvStoreP<S>(d, 1u, c, make_float2(t.px[j], t.py[j]));
    }
}
// The coarse levels from lds_level down, all in one workgroup's LDS:
// a red and a black sweep on each but the coarsest (smoothRed, smoothBlack)
// and restrictLevel into the next from its residuals, the coarsest's SOR
// sweeps (solve), and back up, prolong and the two sweeps again
// (fluid.wgsl:2317-2497), then the levels written back to the field.  The
// first level comes from the field as the pass above left it.  The
// arithmetic is restrictLevel's, smoothAt's, solve's and prolong's, in the
// same order; only where the values wait between passes differs.  A level's
// face betas (faceBeta, :369-374) are taken once, when its rho is known.
//
// The first level is worked by the whole workgroup (its barriers), and so
// are the next while they have more than the world's wave_cells cells; the
// rest by its first wavefront alone, whose lanes pass each phase in step (a
// wavefront's syncs, no barriers), while the other wavefronts wait at the
// next barrier.  Every phase updates its
// cells independently of one another, so how cells are dealt to threads does
// not change a value.
struct LdsLevel { float *px, *py, *rx, *ry, *rho, *bR, *bD; uint32_t w, h; };
// a phase's end: the workgroup's barrier, or the first wavefront's sync
__device__ __forceinline__ void phaseSync(bool wg) {
    if (wg) __syncthreads();
    else waveSync();
}
// A phase's cells of a w-wide level, j = tid, tid + nthr, ..., with their
// column and row stepped along (one division a phase, not one a cell).
struct LinStep { uint32_t x, y, dx, dy, w; };
__device__ __forceinline__ LinStep linStart(uint32_t tid, uint32_t nthr, uint32_t w) {
    LinStep s;
    s.w = w;
    s.x = tid % w;
    s.y = tid / w;
    s.dx = nthr % w;
    s.dy = nthr / w;
    return s;
}
__device__ __forceinline__ void linNext(LinStep &s) {
    s.x += s.dx;
    s.y += s.dy;
    if (s.x >= s.w) {
        s.x -= s.w;
        s.y += 1u;
    }
}
// each cell's right and lower face betas, 0 where the stencil skips
__device__ void ldsBetas(const LdsLevel &L, uint32_t tid, uint32_t nthr) {
    LinStep st = linStart(tid, nthr, L.w);
    for (uint32_t j = tid; j < L.w * L.h; j += nthr, linNext(st)) {
        const uint32_t x = st.x, y = st.y;
        float r = L.rho[j];
        L.bR[j] = x + 1u < L.w ? faceBeta(r, L.rho[j + 1u]) : 0.0f;
        L.bD[j] = y + 1u < L.h ? faceBeta(r, L.rho[j + L.w]) : 0.0f;
    }
}
struct LdsSum { float sx, sy, diag; };
// (as tileStencil; a neighbour off the level is read at the cell itself)
__device__ __forceinline__ LdsSum ldsStencil(const LdsLevel &L, uint32_t x, uint32_t y) {
    uint32_t j = y * L.w + x;
    const float b[4] = {x > 0u ? L.bR[j - 1u] : 0.0f, L.bR[j], y > 0u ? L.bD[j - L.w] : 0.0f, L.bD[j]};
    const uint32_t nb[4] = {x > 0u ? j - 1u : j, x + 1u < L.w ? j + 1u : j, y > 0u ? j - L.w : j, y + 1u < L.h ? j + L.w : j};
    float px[4], py[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        px[k] = L.px[nb[k]];
        py[k] = L.py[nb[k]];
    }
    LdsSum s = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bool use = b[k] != 0.0f;
        s.sx = use ? s.sx + b[k] * px[k] : s.sx;
        s.sy = use ? s.sy + b[k] * py[k] : s.sy;
        s.diag = use ? s.diag + b[k] : s.diag;
    }
    return s;
}
// one half-sweep: the cells of one colour among the first `limit`
__device__ void ldsSmooth(const LdsLevel &L, uint32_t l, int color, float omega, uint32_t limit, uint32_t tid, uint32_t nthr, bool wg) {
    float h2 = (float)(1u << (2u * l));
    uint32_t hw = (L.w + 1u) / 2u;
    LinStep st = linStart(tid, nthr, hw);
    for (uint32_t k = tid; k < hw * L.h; k += nthr, linNext(st)) {
        const uint32_t y = st.y, i = st.x;
        uint32_t x = 2u * i + (((uint32_t)color + y) & 1u), j = y * L.w + x;
        if (x >= L.w || j >= limit) continue;
        float rho = L.rho[j];
        if (rho <= 0.0f) continue;
        LdsSum s = ldsStencil(L, x, y);
        if (s.diag <= 0.0f) continue;
        float sx = (s.sx - h2 * L.rx[j]) / s.diag, sy = (s.sy - h2 * L.ry[j]) / s.diag;
        L.px[j] = wmix(L.px[j], sx, omega);
        L.py[j] = wmix(L.py[j], sy, omega);
    }
    phaseSync(wg);
}
// `pairs` red-black pairs (a red then a black half-sweep) over the cells of
// a level below its first `limit`, by nthr threads: with register sets where
// a colour's cells number at most SWEEP_K a thread, else ldsSmooth's loop
__device__ void ldsSweeps(const LdsLevel &L, uint32_t l, uint32_t pairs, float omega, uint32_t limit, uint32_t tid, uint32_t nthr, bool wg) {
    const uint32_t hw = (L.w + 1u) / 2u;
    if (hw * L.h > SWEEP_K * nthr) {
        for (uint32_t k = 0; k < pairs; k++) {
            ldsSmooth(L, l, 0, omega, limit, tid, nthr, wg);
            ldsSmooth(L, l, 1, omega, limit, tid, nthr, wg);
        }
        return;
    }
    const float h2 = (float)(1u << (2u * l));
    SweepSet red, black;   // (by name: see tileSweeps)
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int m = 0; m < SWEEP_K; m++) {
            SweepSet &set = c == 0 ? red : black;
            sweepNone(set, m);
            const uint32_t k = tid + (uint32_t)m * nthr;
            if (k >= hw * L.h) continue;
            const uint32_t y = k / hw, x = 2u * (k - y * hw) + (((uint32_t)c + y) & 1u), j = y * L.w + x;
            if (x >= L.w || j >= limit || L.rho[j] <= 0.0f) continue;
            sweepCell(set, m, (int32_t)j, x > 0u ? L.bR[j - 1u] : 0.0f, L.bR[j], y > 0u ? L.bD[j - L.w] : 0.0f, L.bD[j],
                      h2, L.rx[j], L.ry[j]);
        }
    sweepVote(red);
    sweepVote(black);
    for (uint32_t s = 0; s < pairs; s++) {
        sweepHalf(red, L.px, L.py, (int32_t)L.w, omega);
        phaseSync(wg);
        sweepHalf(black, L.px, L.py, (int32_t)L.w, omega);
        phaseSync(wg);
    }
}
// restrictLevel (:2342-2377) from level l (T) into level l + 1 (U): each
// child's residual (residualAt, :2306-2315) summed
__device__ void ldsRestrict(const Dev &d, const LdsLevel &T, const LdsLevel &U, uint32_t l, uint32_t tid, uint32_t nthr) {
    const float ih2 = 1.0f / (float)(1u << (2u * l));   // (a power of two: x / h2 and x * ih2 are the same value)
    const int32_t span = 1 << (l + 1u);
    LinStep st = linStart(tid, nthr, U.w);
    for (uint32_t j = tid; j < U.w * U.h; j += nthr, linNext(st)) {
        const uint32_t x = st.x, y = st.y;
        float rho = 0.0f, rxv = 0.0f, ryv = 0.0f;
        bool shortcut = false;
        if (span <= WIG_BLOCK) {
            i2 b = I2((int32_t)x * span / WIG_BLOCK, (int32_t)y * span / WIG_BLOCK);
            shortcut = inBlocks(d, b) && (d.blocks[STATE_SLOT(d, blockIndex(d, b))] & ACTIVE) == 0u;
        }
        if (!shortcut) {
            float fluid = 0.0f, sum = 0.0f, sx = 0.0f, sy = 0.0f;
            bool air = false;
            for (int q = 0; q < 4; q++) {
                uint32_t cx = 2u * x + (uint32_t)(q & 1), cy = 2u * y + (uint32_t)(q >> 1);
                if (cx >= T.w || cy >= T.h) continue;     // outside the level: rho -1
                uint32_t jc = cy * T.w + cx;
                float rk = T.rho[jc];
                if (rk == 0.0f) air = true;
                else if (rk > 0.0f) {
                    fluid += 1.0f;
                    sum += rk;
                    LdsSum st = ldsStencil(T, cx, cy);
                    sx += T.rx[jc] - (st.sx - (st.diag + 0.0f) * T.px[jc]) * ih2;
                    sy += T.ry[jc] - (st.sy - (st.diag + 0.0f) * T.py[jc]) * ih2;
                }
            }
            rho = air ? 0.0f : (fluid > 0.0f ? sum / fluid : -1.0f);
            rxv = rho > 0.0f ? sx * 0.25f : 0.0f;
            ryv = rho > 0.0f ? sy * 0.25f : 0.0f;
        }
        U.px[j] = 0.0f;
        U.py[j] = 0.0f;
        U.rx[j] = rxv;
        U.ry[j] = ryv;
        U.rho[j] = rho;
    }
}
// prolong (:2459-2497) from level l + 1 (U) into level l (T)
__device__ void ldsProlong(const LdsLevel &T, const LdsLevel &U, uint32_t tid, uint32_t nthr) {
    LinStep st = linStart(tid, nthr, T.w);
    for (uint32_t j = tid; j < T.w * T.h; j += nthr, linNext(st)) {
        const uint32_t x = st.x, y = st.y;
        if (T.rho[j] <= 0.0f) continue;
        i2 coarse = I2((int32_t)x / 2, (int32_t)y / 2);
        i2 sg = I2((x & 1u) == 1u ? 1 : -1, (y & 1u) == 1u ? 1 : -1);
        float rc[4], cx[4], cy[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            i2 o = k == 0 ? I2(0, 0) : (k == 1 ? I2(sg.x, 0) : (k == 2 ? I2(0, sg.y) : sg));
            i2 at = I2(coarse.x + o.x, coarse.y + o.y);
            const bool in = !(at.x < 0 || at.y < 0 || at.x >= (int32_t)U.w || at.y >= (int32_t)U.h);   // else rho -1
            const uint32_t ja = in ? (uint32_t)at.y * U.w + (uint32_t)at.x : 0u;
            const float r = U.rho[ja];
            rc[k] = in ? r : -1.0f;
            cx[k] = U.px[ja];
            cy[k] = U.py[ja];
        }
        const ProlongSum ps = prolongSum(rc, cx, cy);
        float dn = fmaxf(ps.weight, 1e-6f);
        T.px[j] = T.px[j] + ps.sx / dn;
        T.py[j] = T.py[j] + ps.sy / dn;
    }
}
template <int S> __global__ void __launch_bounds__(1024) k_coarseLDS(WorldGrid g) {
    WORLD_REF(g);
    if (d.store != S) return;
    const Iter it = iterWords(d);
    extern __shared__ float sm[];
    uint32_t l0 = d.lds_level;
    const uint32_t L = d.L;
    const uint32_t tid = threadIdx.x, W = (uint32_t)d.W, H = (uint32_t)d.H;
    const uint32_t pairs = d.level_pairs, csweeps = d.coarse_sweeps;
    // level l's arrays in LDS, levels l0 .. L one after another (the level
    // sizes as the world's: side rounded up at each halving)
    auto level = [&](uint32_t l) {
        uint32_t off = 0;
        for (uint32_t k = l0; k < l; k++) off += LDS_LEVEL_FLOATS * ((W + (1u << k) - 1u) >> k) * ((H + (1u << k) - 1u) >> k);
        LdsLevel T;
        T.w = (W + (1u << l) - 1u) >> l;
        T.h = (H + (1u << l) - 1u) >> l;
        uint32_t n = T.w * T.h;
        T.px = sm + off;
        T.py = sm + off + n;
        T.rx = sm + off + 2u * n;
        T.ry = sm + off + 3u * n;
        T.rho = sm + off + 4u * n;
        T.bR = sm + off + 5u * n;
        T.bD = sm + off + 6u * n;
        return T;
    };
    // level l0 as the pass above left it in the field (or, with no blocks
    // worked on, the coarsest as it was)
    auto loadFirst = [&]() {
        // (every load made at once: six a thread hold what LDS can)
        const LdsLevel T = level(l0);
        const uint32_t n = T.w * T.h;
        float4 e[6];
        float r[6];
#pragma unroll
        for (uint32_t q = 0; q < 6u; q++) {
            if (q * 1024u >= n) break;
            const uint32_t j = min(tid + q * 1024u, n - 1u);
            const i2 c = I2((int32_t)(j % T.w), (int32_t)(j / T.w));
            e[q] = lvLoad<S>(d, l0, c);
            r[q] = lvRho<S>(d, l0, c);
        }
#pragma unroll
        for (uint32_t q = 0; q < 6u; q++) {
            const uint32_t j = tid + q * 1024u;
            if (q * 1024u >= n) break;
            if (j >= n) continue;
            T.px[j] = e[q].x;
            T.py[j] = e[q].y;
            T.rx[j] = e[q].z;
            T.ry[j] = e[q].w;
            T.rho[j] = r[q];
        }
        for (uint32_t j = tid + 6u * 1024u; j < n; j += 1024u) {   // (none while LDS holds 160 KB)
            const i2 c = I2((int32_t)(j % T.w), (int32_t)(j / T.w));
            const float4 ee = lvLoad<S>(d, l0, c);
            T.px[j] = ee.x;
            T.py[j] = ee.y;
            T.rx[j] = ee.z;
            T.ry[j] = ee.w;
            T.rho[j] = lvRho<S>(d, l0, c);
        }
    };
    if (!it.encoded || it.converged) return;
    const bool idle = it.active == 0u;   // restrict and prolong pass over nothing; solve still sweeps
    if (idle) l0 = L;
    loadFirst();
    __syncthreads();
    ldsBetas(level(l0), tid, 1024u);
    __syncthreads();
    // The levels from lw down are the first wavefront's: the first below
    // l0 with at most wave_cells cells (L + 1: none).
    uint32_t lw = L + 1u;
    for (uint32_t l = L; l > l0 && ((W + (1u << l) - 1u) >> l) * ((H + (1u << l) - 1u) >> l) <= d.wave_cells; l--) lw = l;
    // down: the sweeps, then restrictLevel into the next level
    for (uint32_t l = l0; l < L && l < lw && !idle; l++) {
        const LdsLevel T = level(l), U = level(l + 1u);
        ldsSweeps(T, l, pairs, 1.0f, 0xffffffffu, tid, 1024u, true);
        ldsRestrict(d, T, U, l, tid, 1024u);
        __syncthreads();
        ldsBetas(U, tid, 1024u);
        __syncthreads();
    }
    if (lw > L) {
        // the coarsest (solve, :2407-2456), by the workgroup: its first
        // MG_COARSEST cells
        const LdsLevel T = level(L);
        ldsSweeps(T, L, csweeps, OMEGA, (uint32_t)MG_COARSEST, tid, 1024u, true);
    } else if (tid < 64u) {
        // the first wavefront's levels: down, the coarsest, up to lw
        for (uint32_t l = lw; l < L && !idle; l++) {
            const LdsLevel T = level(l), U = level(l + 1u);
            ldsSweeps(T, l, pairs, 1.0f, 0xffffffffu, tid, 64u, false);
            ldsRestrict(d, T, U, l, tid, 64u);
            phaseSync(false);
            ldsBetas(U, tid, 64u);
            phaseSync(false);
        }
        {
            const LdsLevel T = level(L);
            ldsSweeps(T, L, csweeps, OMEGA, (uint32_t)MG_COARSEST, tid, 64u, false);
        }
        // up: prolong (:2459-2497), then a red and a black sweep
        for (int32_t l = (int32_t)L - 1; l >= (int32_t)lw && !idle; l--) {
            const LdsLevel T = level((uint32_t)l), U = level((uint32_t)l + 1u);
            ldsProlong(T, U, tid, 64u);
            phaseSync(false);
            ldsSweeps(T, (uint32_t)l, pairs, 1.0f, 0xffffffffu, tid, 64u, false);
        }
    }
    __syncthreads();
    // up the workgroup's levels
    for (int32_t l = (int32_t)(lw < L ? lw : L) - 1; l >= (int32_t)l0 && !idle; l--) {
        const LdsLevel T = level((uint32_t)l), U = level((uint32_t)l + 1u);
        ldsProlong(T, U, tid, 1024u);
        __syncthreads();
        ldsSweeps(T, (uint32_t)l, pairs, 1.0f, 0xffffffffu, tid, 1024u, true);
    }
    // the levels back into the field
    for (uint32_t l = l0; l <= L; l++) {
        const LdsLevel T = level(l);
        LinStep st = linStart(tid, 1024u, T.w);
        for (uint32_t j = tid; j < T.w * T.h; j += 1024u, linNext(st)) {
            const i2 c = I2((int32_t)st.x, (int32_t)st.y);
            lvStore<S>(d, l, c, make_float4(T.px[j], T.py[j], T.rx[j], T.ry[j]));
            if (!idle) lvStoreRho<S>(d, l, c, T.rho[j]);
        }
    }
}

// The step's Params words into device memory (k_setStep), before the step.

// ------------------------------------------------ conjugate gradients
// cgDirection's p (:2767-2776) at cell i: z + beta p, 0 off the fluid
template <int S> __device__ __forceinline__ float newP(const Dev &d, uint32_t i, double beta, const float *pold) {
    const uint32_t f = d.cflag[i];
    const float z = fnLoad<S>(d, 0, i).x, po = pold[i];   // (loaded with the flag, not after it)
    return (f & CF_FLUID) ? (float)((double)z + beta * (double)po) : 0.0f;
}
// applyA (:2595-2613) from the packed betas: every face but an obstacle's
// counts in the diagonal, only a fluid neighbour's p in the sum (nb: the
// neighbours' new p, beside()'s order)
__device__ __forceinline__ float applyA(float4 cf, uint32_t f, float p, const float *nb) {
    const float v[4] = {cf.x, cf.y, cf.z, cf.w};
    float sum = 0.0f, diag = 0.0f;
    for (int k = 0; k < 4; k++) {
        float b = fabsf(v[k]);
        if (b == 0.0f) continue;
        if (__float_as_uint(v[k]) >> 31) sum += b * nb[k];
        diag += b;
    }
    diag += (f & CF_PINNED) ? PINF : 0.0f;
    return diag * p - sum;
}
__device__ __forceinline__ float *pBuf(const Dev &d, uint32_t k) { return (k & 1u) ? d.cp1 : d.cp0; }

__global__ void __launch_bounds__(64) k_cgStart(WorldGrid g) {                     // :2665-2674
    WORLD(g);
    if (!d.full) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (!inWorld(d, c)) continue;
        uint32_t i = cellIndex(d, c), pi = packIndex(d, c);
        d.cx[pi] = gridX(d, c);
        d.cr[pi] = 0.0f;
        d.cq[pi] = 0.0f;
        d.cp0[pi] = 0.0f;
        CNT(d, i, G_REGION) = i;
    }
}
__global__ void __launch_bounds__(64) k_cgAnchor(WorldGrid g) {                    // :2692-2705
    WORLD(g);
    if (!d.full) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (!inWorld(d, c) || gridZ(d, c) <= 0.0f) continue;
        for (int k = 0; k < 4; k++) {
            i2 n = beside(c, k);
            if (inWorld(d, n) && rhoAt(d, 0u, n) == 0.0f) {
                uint32_t r = ufFind(d, cellIndex(d, c));
                if ((ufLink(d, r) & REGION_OPEN) == 0u) atomicOr(&CNT(d, r, G_REGION), REGION_OPEN);
                break;
            }
        }
    }
}
// ------------------------------------------------ FP64 sums
// A pass's block partials are kept by the block's place in the active list
// (which is in block-index order) and summed by one workgroup of 1024, lane
// i taking places i, i + 1024, ..., then halving: the same order every run.
// (The sum is a launch of its own.  A pass whose last workgroup took it needs
// every workgroup to fence its partials out of its XCD's L2, a write-back
// each: on MI355X that made cgRho, cgApply and cgUpdate five to ten times
// slower than the one-workgroup launch costs.)

__device__ __forceinline__ double ld_f64(const double *p) {
    unsigned long long u = __hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)u);
}
// (1024 lanes, lane i taking places i, i + 1024, ..., then halving; two sums
// at once where a pass has two)
#define SUM_LANES 1024u
// Each lane's places, SUM_HELD of them (16384 places in all), held in
// registers with their loads issued at once (those below the count: one
// workgroup's loads are the launch's, and places past the count would
// only slow it); places past those are read as the sum goes.
#define SUM_HELD 16
template <int F> struct PlaceHold { double v[F][SUM_HELD]; uint32_t na; };
template <int F> __device__ __forceinline__ void holdPlaces(const Dev &d, uint32_t na, const uint32_t (&w)[F], PlaceHold<F> &h) {
    h.na = na;
#pragma unroll
    for (int q = 0; q < SUM_HELD; q++) {
        const uint32_t j = threadIdx.x + (uint32_t)q * SUM_LANES;
#pragma unroll
        for (int f = 0; f < F; f++) h.v[f][q] = j < na ? d.bpart[4ull * j + w[f]] : 0.0;
    }
}
// this lane's sum of word w (held as field f): places lane, lane +
// SUM_LANES, ... below the count, in that order
template <int F> __device__ __forceinline__ double laneSum(const Dev &d, const PlaceHold<F> &h, int f, uint32_t w) {
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < SUM_HELD; q++)
        if (threadIdx.x + (uint32_t)q * SUM_LANES < h.na) a += h.v[f][q];
    for (uint32_t j = threadIdx.x + SUM_HELD * SUM_LANES; j < h.na; j += SUM_LANES) a += d.bpart[4ull * j + w];
    return a;
}
// The lanes' values halved, lane i taking lane i + h for h = 512 .. 1:
// through LDS while h spans wavefronts, then within the first wavefront
// (its shuffles take the same operands in the same order); the totals to
// every thread.  m: a max over the lanes alongside, by the same halving.
__device__ __forceinline__ void laneTree(double a0, double a1, uint32_t m, double *s0, double *s1, uint32_t *sm,
                                         double *o0, double *o1, uint32_t *om) {
    const uint32_t li = threadIdx.x;
    s0[li] = a0;
    s1[li] = a1;
    if (sm) sm[li] = m;
    __syncthreads();
    for (uint32_t h = SUM_LANES / 2u; h >= 64u; h >>= 1) {
        if (li < h) {
            s0[li] += s0[li + h];
            s1[li] += s1[li + h];
            if (sm) sm[li] = max(sm[li], sm[li + h]);
        }
        __syncthreads();
    }
    if (li < 64u) {
        double v0 = wave_tree_sum(s0[li]), v1 = wave_tree_sum(s1[li]);
        uint32_t vm = sm ? wave_max_u32(sm[li]) : 0u;
        if (li == 0u) {
            s0[0] = v0;
            s1[0] = v1;
            if (sm) sm[0] = vm;
        }
    }
    __syncthreads();
    *o0 = s0[0];
    *o1 = s1[0];
    if (sm) *om = sm[0];
    __syncthreads();
}
#define FOR_ACTIVE_BLOCKS4(d, wid) \
    for (uint32_t wid = bid * WAVES + threadIdx.x / 64u, n_ = activeCount(d); wid < n_; wid += nbid * WAVES)
#define LANE (threadIdx.x % 64u)
// The iteration's passes over the fine level take NB places of the active
// list a wavefront at once, so that NB blocks' loads are in flight together
// (each block's sums are still its own wavefront halving, as above).
#define NB 4u
// this lane's cell in the blocks at places w0 .. w0 + NB - 1, as packed
// indices (past the list's end block 0's: loads there are harmless, stores
// are not made).  Every lane of a block has its slot; those outside the
// world hold cflag 0.  A pass reads its first places' list entries beside
// the iteration's words (the loads clamped into the list, made whatever the
// count), and each next places' while it works on these.
#define PLACES_FIRST ((bid * WAVES + (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x / 64u)) * NB)
#define PLACES_STRIDE (nbid * WAVES * NB)
__device__ __forceinline__ void placeEntries(const Dev &d, uint32_t w0, uint32_t *e) {
    for (uint32_t j = 0; j < NB; j++) e[j] = d.blocks[ACTIVE_SLOT(d, min(w0 + j, d.B - 1u))];
}
__device__ __forceinline__ void placeCellsFrom(const Iter &it, uint32_t w0, const uint32_t *e, uint32_t *i) {
    for (uint32_t j = 0; j < NB; j++) i[j] = (w0 + j < it.active ? e[j] : 0u) * 64u + LANE;
}
// the packed index of the cell beside packed cell i (beside()'s k) in the
// next block over, for a lane on that edge of its block; i itself where
// that is off the packed arrays (a cell outside the world, whose beta is 0)
// the cell of packed index i
__device__ __forceinline__ i2 packedCell(const Dev &d, uint32_t i) {
    const uint32_t b = i / 64u, l = i % 64u;
    return I2((int32_t)((b % d.bw) * 8u + (l & 7u)), (int32_t)((b / d.bw) * 8u + (l >> 3)));
}
__device__ __forceinline__ uint32_t acrossEdge(const Dev &d, uint32_t i, int k) {
    uint32_t n = k == 0 ? i - 57u : (k == 1 ? i + 57u : (k == 2 ? i - 64u * d.bw + 56u : i + 64u * d.bw - 56u));
    return n < 64u * d.B ? n : i;
}

// cgResidual (:2708-2717), and ||b||^2 for the relative stop
__global__ void __launch_bounds__(256) k_cgResidual(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    if (bid == 0 && threadIdx.x == 0) {
        SolveHead h = {};
        *d.head = h;
        // the first convergence checkpoint: WIG_CHUNK_AUTO takes as many
        // iterations as the last solve that stopped at one ran
        uint32_t last = d.wstate[WS_LAST];
        d.wstate[WS_FIRST] = d.first_chunk == 0xffffffffu ? (last > 1u ? last : 1u) : d.first_chunk;
        d.wstate[WS_STOP] = NO_STOP;
        d.wstate[WS_CYCLE] = 0u;
    }
    FOR_ACTIVE_BLOCKS4(d, wid) {
        i2 c = blockCell(activeBlock(d, wid), LANE);
        double bb = 0.0;
        uint32_t most = 0;
        if (inWorld(d, c) && gridZ(d, c) > 0.0f) {
            uint32_t i = cellIndex(d, c), pi = packIndex(d, c);
            float r = -residualAt(d, 0u, c).x;
            if (d.body_on) r = r - couplingAt(d, c, PUSH_X);   // (:2714-2716)
            d.cr[pi] = r;
            most = ubits(fabsf(r));
            double bi = -(double)FLD(d, 2u * i, 1);
            bb = bi * bi;
            if (d.capb) d.capb[pi] = -FLD(d, 2u * i, 1);
            if (d.replace_every) d.cb[pi] = -FLD(d, 2u * i, 1);
        }
        bb = wave_tree_sum(bb);
        most = wave_max_u32(most);
        if (LANE == 0) {
            d.bpart[4ull * wid + 2] = bb;
            d.bpart[4ull * wid + 3] = (double)most;
        }
    }
}
__global__ void __launch_bounds__(1024) k_sumNorm(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    __shared__ double s0[SUM_LANES], s1[SUM_LANES];
    __shared__ uint32_t sMost[SUM_LANES];
    // ||b||^2, and the largest |r0| (a block's as f32 bits), for the 16-bit
    // V-cycle's scale: one halving, as k_cgCheck's
    PlaceHold<2> ph;
    const uint32_t w[2] = {2u, 3u};
    holdPlaces<2>(d, activeCount(d), w, ph);
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < SUM_HELD; q++) {
        const uint32_t b = (uint32_t)ph.v[1][q];
        if (threadIdx.x + (uint32_t)q * SUM_LANES < ph.na) m = b > m ? b : m;
    }
    for (uint32_t j = threadIdx.x + SUM_HELD * SUM_LANES; j < ph.na; j += SUM_LANES) {
        uint32_t b = (uint32_t)d.bpart[4ull * j + 3];
        m = b > m ? b : m;
    }
    double bb, unused;
    uint32_t most;
    laneTree(laneSum<2>(d, ph, 0, 2u), 0.0, m, s0, s1, sMost, &bb, &unused, &most);
    if (threadIdx.x == 0) {
        d.head->bb = bb;
        d.head->vscale = vcycleScale(d, most);
    }
}
// cgPrepare (:2721-2732), and the solve's fine level packed (the arrays'
// layout is with the Dev struct).  p is 0 and the drift as it stands, the
// rhs -r and the drift rhs; the ring of cells beside the blocks worked on
// holds its field values, as the relaxations' rims read them there.
template <int S> __global__ void __launch_bounds__(64) k_cgPrepare(WorldGrid g) {
    WORLD(g);
    if (d.store != S) return;
    if (!d.full) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 bk = activeBlockAt(d, blk_), c = blockCell(bk, threadIdx.x);
        if (!inWorld(d, c)) continue;
        uint32_t i = cellIndex(d, c), pi = packIndex(d, c);
        float rho = FLD(d, 2u * i, 2);
        float2 v = make_float2(0.0f, FLD(d, 2u * i + 1u, 0));
        fnStore<S>(d, 0, pi, v);
        fnStore<S>(d, 1, pi, v);
        uint32_t f = 0;
        float4 cf = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float2 r = make_float2(0.0f, 0.0f);
        if (rho > 0.0f) {
            f = CF_FLUID | (pinnedAt(d, i) ? CF_PINNED : 0u);
            if (d.body_on)
                for (int k = 0; k < 4; k++)
                    if (rigidAt(d, beside(c, k)) != WNONE) f |= CF_BODY;
            float dc = densityAt(d, c), b[4];
            for (int k = 0; k < 4; k++) {
                i2 n = beside(c, k);
                float rn = gridZ(d, n), beta = fineBeta(rho, rn, dc, densityAt(d, n));
                b[k] = rn > 0.0f ? -beta : beta;
            }
            cf = make_float4(b[0], b[1], b[2], b[3]);
            r = make_float2(S == STORE_F32 ? -d.cr[pi] : -d.cr[pi] * d.head->vscale, FLD(d, 2u * i + 1u, 1));
        }
        d.coef[pi] = cf;
        d.cflag[pi] = (uint8_t)f;
        fnStore<S>(d, 2, pi, r);
        for (int k = 0; k < 4; k++) {
            i2 n = beside(c, k);
            if (!inWorld(d, n) || (idiv8(n).x == bk.x && idiv8(n).y == bk.y) || isActiveBlock(d, idiv8(n))) continue;
            uint32_t ni = cellIndex(d, n), pn = packIndex(d, n);
            float2 q = make_float2(FLD(d, 2u * ni, 0), FLD(d, 2u * ni + 1u, 0));
            fnStore<S>(d, 0, pn, q);
            fnStore<S>(d, 1, pn, q);
            d.cflag[pn] = FLD(d, 2u * ni, 2) > 0.0f ? (uint8_t)CF_FLUID : (uint8_t)0;
        }
    }
}
// cgRho + cgBeta (:2736-2764): r.z and z.Ap, and Notay's flexible beta in FP64
template <int S> __global__ void __launch_bounds__(256) k_cgRho(WorldGrid g) {
    WORLD(g);
    if (d.store != S) return;
    const uint32_t wf = PLACES_FIRST;
    uint32_t e[NB];
    placeEntries(d, wf, e);
    const Iter it = iterWords(d);
    if (!it.encoded || it.converged) return;
    for (uint32_t w0 = wf; w0 < it.active; w0 += PLACES_STRIDE) {
        uint32_t i[NB];
        placeCellsFrom(it, w0, e, i);
        placeEntries(d, w0 + PLACES_STRIDE, e);
        uint32_t f[NB];
        float z[NB], r[NB], q[NB];
        for (uint32_t j = 0; j < NB; j++) {
            f[j] = d.cflag[i[j]];
            z[j] = fnLoad<S>(d, 0, i[j]).x;
            r[j] = d.cr[i[j]];
            q[j] = d.cq[i[j]];
        }
        for (uint32_t j = 0; j < NB; j++) {
            double v0 = 0.0, v1 = 0.0;
            if (f[j] & CF_FLUID) {
                v0 = (double)r[j] * (double)z[j];
                v1 = (double)q[j] * (double)z[j];
            }
            v0 = wave_tree_sum(v0);
            v1 = wave_tree_sum(v1);
            if (LANE == 0 && w0 + j < it.active) {
                d.bpart[4ull * (w0 + j) + 0] = v0;
                d.bpart[4ull * (w0 + j) + 1] = v1;
            }
        }
    }
}
__global__ void __launch_bounds__(1024) k_cgBeta(WorldGrid g) {
    WORLD(g);
    const Iter it = iterWords(d);
    PlaceHold<2> ph;
    const uint32_t w[2] = {0u, 1u};
    holdPlaces<2>(d, it.active, w, ph);
    const uint32_t cycle = it.cycle;
    if (!it.encoded || it.converged) return;
    __shared__ double t0[SUM_LANES], t1[SUM_LANES];
    double s0, s1;
    laneTree(laneSum<2>(d, ph, 0, 0u), laneSum<2>(d, ph, 1, 1u), 0u, t0, t1, nullptr, &s0, &s1, nullptr);
    if (threadIdx.x == 0) {
        SolveHead *h = d.head;
        double beta = (cycle == 0u || h->pq <= 0.0) ? 0.0 : -s1 / h->pq;
        h->rz = s0;
        h->beta = beta;
        CGV(d, d.C + d.B, 0) = (float)s0;
        CGV(d, d.C + d.B, 2) = (float)beta;
    }
}
// cgDirection + cgApply + cgAlpha (:2767-2814): the iteration's p (into
// the other p buffer) and q = A p
template <int S> __global__ void __launch_bounds__(256) k_cgApply(WorldGrid g) {
    WORLD(g);
    if (d.store != S) return;
    const uint32_t wf = PLACES_FIRST;
    uint32_t e[NB];
    placeEntries(d, wf, e);
    const Iter it = iterWords(d);
    const double beta = ld_f64(&d.head->beta);
    if (!it.encoded || it.converged) return;
    const float *pold = pBuf(d, it.cycle);
    float *pnew = pBuf(d, it.cycle + 1u);
    const uint32_t x = LANE & 7u, y = LANE >> 3;
    const bool edge[4] = {x == 0u, x == 7u, y == 0u, y == 7u};
    const uint32_t from[4] = {(LANE + 63u) & 63u, (LANE + 1u) & 63u, (LANE + 56u) & 63u, (LANE + 8u) & 63u};
    // the cells beside the block's four edges, one a lane of lanes 0 .. 31
    // (left, right, upper, lower edge, 8 each), and the lane a cell on an
    // edge finds its neighbour's p at
    const uint32_t ek = (LANE / 8u) & 3u, em = LANE % 8u;
    const uint32_t eat = ek == 0u ? em * 8u : (ek == 1u ? em * 8u + 7u : (ek == 2u ? em : 56u + em));
    const uint32_t slot[4] = {y, 8u + y, 16u + x, 24u + x};
    for (uint32_t w0 = wf; w0 < it.active; w0 += PLACES_STRIDE) {
        uint32_t i[NB];
        placeCellsFrom(it, w0, e, i);
        placeEntries(d, w0 + PLACES_STRIDE, e);
        uint32_t f[NB];
        float z[NB], po[NB], pe[NB];
        float4 cf[NB];
        for (uint32_t j = 0; j < NB; j++) {
            f[j] = d.cflag[i[j]];
            z[j] = fnLoad<S>(d, 0, i[j]).x;
            po[j] = pold[i[j]];
            cf[j] = d.coef[i[j]];
        }
        // newP beside the edges (lanes 32 .. 63 take their own cell's, unused)
        for (uint32_t j = 0; j < NB; j++) {
            const uint32_t b0 = i[j] - LANE;
            pe[j] = newP<S>(d, LANE < 32u ? acrossEdge(d, b0 + eat, (int)ek) : i[j], beta, pold);
        }
        for (uint32_t j = 0; j < NB; j++) {
            // cgDirection's p (newP) at this cell, the neighbours' inside the
            // block from their lanes, across its edges from the edge lanes
            float p = (f[j] & CF_FLUID) ? (float)((double)z[j] + beta * (double)po[j]) : 0.0f;
            float nb[4];
            for (int k = 0; k < 4; k++) {
                float o = __shfl(p, (int)from[k], WAVE);
                float e = __shfl(pe[j], (int)slot[k], WAVE);
                nb[k] = edge[k] ? e : o;
            }
            float q = (f[j] & CF_FLUID) ? applyA(cf[j], f[j], p, nb) : 0.0f;
            // the bodies' give, J^T M^-1 J p (:2790)
            if (d.body_on && (f[j] & CF_FLUID))
                q = q + ((f[j] & CF_BODY) ? couplingAt(d, packedCell(d, i[j]), PUSH_P0 + it.cycle % 2u) : 0.0f);
            double v = (double)p * (double)q;
            v = wave_tree_sum(v);
            if (w0 + j < it.active) {
                pnew[i[j]] = p;
                d.cq[i[j]] = q;
                if (LANE == 0) d.bpart[4ull * (w0 + j) + 0] = v;
            }
        }
    }
}
__global__ void __launch_bounds__(1024) k_cgAlpha(WorldGrid g) {
    WORLD(g);
    const Iter it = iterWords(d);
    PlaceHold<1> ph;
    const uint32_t w[1] = {0u};
    holdPlaces<1>(d, it.active, w, ph);
    if (!it.encoded || it.converged) return;
    __shared__ double t0[SUM_LANES], t1[SUM_LANES];
    double pq, unused;
    laneTree(laneSum<1>(d, ph, 0, 0u), 0.0, 0u, t0, t1, nullptr, &pq, &unused, nullptr);
    if (threadIdx.x == 0) {
        SolveHead *h = d.head;
        h->alpha = pq <= 0.0 ? 0.0 : h->rz / pq;
        h->pq = pq;
        CGV(d, d.C + d.B, 1) = (float)h->alpha;
        CGV(d, d.C + d.B, 3) = (float)pq;
    }
    // the bodies' sums of this iteration's direction, done with, emptied for
    // the iteration after next (:2808-2813)
    if (d.body_on)
        for (uint32_t k = threadIdx.x; k < MAX_BODIES * 6u; k += blockDim.x)
            d.pushes[pushSlot(k / 6u, PUSH_P0 + it.cycle % 2u, 0u) + k % 6u] = 0u;
}
// cgUpdate (:2852-2880), and whether the next
// iteration runs: max |r| under the tolerance (the reference's converged(),
// :2223-2226), or ||r|| <= tol ||b||
template <int S> __global__ void __launch_bounds__(256) k_cgUpdate(WorldGrid g) {
    WORLD(g);
    if (d.store != S) return;
    const uint32_t wf = PLACES_FIRST;
    uint32_t e[NB];
    placeEntries(d, wf, e);
    const Iter it = iterWords(d);
    const double alpha = ld_f64(&d.head->alpha);
    const float vscale = d.head->vscale;
    if (!it.encoded) return;
    const uint32_t cycle = it.cycle;
    if (bid == 0 && threadIdx.x == 0) {
        d.blocks[COUNT_SLOT(d, COUNT_RESIDUAL + (cycle + 1u) % 3u)] = 0u;
        if (!it.converged) atomicAdd(&d.blocks[COUNT_SLOT(d, COUNT_CYCLES)], 1u);
    }
    if (it.converged) return;
    const float *pn = pBuf(d, cycle + 1u);
    for (uint32_t w0 = wf; w0 < it.active; w0 += PLACES_STRIDE) {
        uint32_t i[NB];
        placeCellsFrom(it, w0, e, i);
        placeEntries(d, w0 + PLACES_STRIDE, e);
        uint32_t f[NB];
        float x[NB], r[NB], p[NB], q[NB];
        for (uint32_t j = 0; j < NB; j++) {
            f[j] = d.cflag[i[j]];
            x[j] = d.cx[i[j]];
            r[j] = d.cr[i[j]];
            p[j] = pn[i[j]];
            q[j] = d.cq[i[j]];
        }
        for (uint32_t j = 0; j < NB; j++) {
            const bool mine = w0 + j < it.active;
            double rr = 0.0;
            uint32_t most = 0;
            if (mine && (f[j] & CF_FLUID)) {
                float rn = (float)((double)r[j] - alpha * (double)q[j]);
                float xn = (float)((double)x[j] + alpha * (double)p[j]);
                d.cx[i[j]] = xn;
                d.cr[i[j]] = rn;
                fnStoreX<S>(d, 0, i[j], 0.0f);
                fnStoreX<S>(d, 2, i[j], S == STORE_F32 ? -rn : -rn * vscale);
                most = ubits(fabsf(rn));
                rr = (double)rn * (double)rn;
            }
            rr = wave_tree_sum(rr);
            most = wave_max_u32(most);
            if (mine && LANE == 0) {
                d.bpart[4ull * (w0 + j) + 2] = rr;
                d.bpart[4ull * (w0 + j) + 3] = (double)most;   // (the block's max |r|, as f32 bits; cgCheck takes the most)
            }
        }
    }
}
// Residual replacement (WigStepCtl.residual_every, this library's): after
// every replace_every-th iteration, r = b - A x from the updated x, the
// products summed in FP64 in beside()'s order, in place of the recursive
// r; with it the V-cycle's next rhs, and the blocks' r.r and max |r| that
// cgCheck judges the stop by.  Beside them each block's sum of
// (|A| |x|)^2: an f32 iterate's rounding leaves ||b - A x|| near
// 2^-24 ||(|A| |x|)|| / 4, the most a solve in f32 can attain.
template <int S> __global__ void __launch_bounds__(256) k_cgTrue(WorldGrid g) {
    WORLD(g);
    if (d.store != S || d.replace_every == 0u) return;
    const uint32_t wf = PLACES_FIRST;
    uint32_t e[NB];
    placeEntries(d, wf, e);
    const Iter it = iterWords(d);
    const float vscale = d.head->vscale;
    if (!it.encoded || it.converged || (it.cycle + 1u) % d.replace_every != 0u) return;
    const uint32_t x = LANE & 7u, y = LANE >> 3;
    for (uint32_t w0 = wf; w0 < it.active; w0 += PLACES_STRIDE) {
        uint32_t i[NB];
        placeCellsFrom(it, w0, e, i);
        placeEntries(d, w0 + PLACES_STRIDE, e);
        uint32_t f[NB];
        float4 cf[NB];
        float xs[NB], bs[NB], xn[NB][4];
        for (uint32_t j = 0; j < NB; j++) {
            f[j] = d.cflag[i[j]];
            cf[j] = d.coef[i[j]];
            xs[j] = d.cx[i[j]];
            bs[j] = d.cb[i[j]];
            xn[j][0] = d.cx[x == 0u ? acrossEdge(d, i[j], 0) : i[j] - 1u];
            xn[j][1] = d.cx[x == 7u ? acrossEdge(d, i[j], 1) : i[j] + 1u];
            xn[j][2] = d.cx[y == 0u ? acrossEdge(d, i[j], 2) : i[j] - 8u];
            xn[j][3] = d.cx[y == 7u ? acrossEdge(d, i[j], 3) : i[j] + 8u];
        }
        for (uint32_t j = 0; j < NB; j++) {
            const bool mine = w0 + j < it.active;
            double rr = 0.0, aa = 0.0;
            uint32_t most = 0;
            if (mine && (f[j] & CF_FLUID)) {
                const float v[4] = {cf[j].x, cf[j].y, cf[j].z, cf[j].w};
                double sum = 0.0, diag = 0.0, mag = 0.0;
                for (int k = 0; k < 4; k++) {
                    const float b = fabsf(v[k]);
                    if (b == 0.0f) continue;
                    if (__float_as_uint(v[k]) >> 31) {
                        sum += (double)b * (double)xn[j][k];
                        mag += (double)b * fabs((double)xn[j][k]);
                    }
                    diag += (double)b;
                }
                if (f[j] & CF_PINNED) diag += (double)PINF;
                const float rn = (float)((double)bs[j] - (diag * (double)xs[j] - sum));
                d.cr[i[j]] = rn;
                fnStoreX<S>(d, 2, i[j], S == STORE_F32 ? -rn : -rn * vscale);
                most = ubits(fabsf(rn));
                rr = (double)rn * (double)rn;
                const double a = diag * fabs((double)xs[j]) + mag;
                aa = a * a;
            }
            rr = wave_tree_sum(rr);
            aa = wave_tree_sum(aa);
            most = wave_max_u32(most);
            if (mine && LANE == 0) {
                d.bpart[4ull * (w0 + j) + 1] = aa;
                d.bpart[4ull * (w0 + j) + 2] = rr;
                d.bpart[4ull * (w0 + j) + 3] = (double)most;
            }
        }
    }
}
// (and the iteration counter stepped, the iteration's last pass: the
// counter lives in device memory so a captured iteration replays alike)
__device__ __forceinline__ void stepCycle(const Dev &d, uint32_t cycle) {
    if (threadIdx.x == 0) __hip_atomic_store(&d.wstate[WS_CYCLE], cycle + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__global__ void __launch_bounds__(1024) k_cgCheck(WorldGrid g) {
    WORLD(g);
    const Iter it = iterWords(d);
    PlaceHold<2> ph;
    const uint32_t w[2] = {2u, 3u};
    holdPlaces<2>(d, it.active, w, ph);
    const uint32_t cycle = it.cycle;
    if (!d.full) return;
    if (!it.encoded || it.converged) {
        stepCycle(d, cycle);
        return;
    }
    __shared__ double t0[SUM_LANES], t1[SUM_LANES];
    __shared__ uint32_t sMost[SUM_LANES];
    uint32_t *slot = &d.blocks[COUNT_SLOT(d, COUNT_RESIDUAL + cycle % 3u)];
    // the sum of r.r, and the most |r| left in any cell (a block's as f32
    // bits; non-negative floats order as their bits)
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < SUM_HELD; q++) {
        const uint32_t b = (uint32_t)ph.v[1][q];
        if (threadIdx.x + (uint32_t)q * SUM_LANES < ph.na) m = b > m ? b : m;
    }
    for (uint32_t j = threadIdx.x + SUM_HELD * SUM_LANES; j < ph.na; j += SUM_LANES) {
        uint32_t b = (uint32_t)d.bpart[4ull * j + 3];
        m = b > m ? b : m;
    }
    // (and, at a replaced residual, sum (|A| |x|)^2, cgTrue's: places in
    // the order laneSum takes them)
    const bool replaced = d.replace_every != 0u && (cycle + 1u) % d.replace_every == 0u;
    double aLane = 0.0;
    if (replaced)
        for (uint32_t j = threadIdx.x; j < ph.na; j += SUM_LANES) aLane += d.bpart[4ull * j + 1];
    double rr, aa;
    uint32_t most;
    laneTree(laneSum<2>(d, ph, 0, 2u), aLane, m, t0, t1, sMost, &rr, &aa, &most);
    if (threadIdx.x == 0) {
        if (most > *slot) *slot = most;
        d.head->vscale = vcycleScale(d, most);
        SolveHead *h = d.head;
        h->rr = rr;
        h->cycles += 1u;
        bool c2;
        if (d.tol_mode == 1u) {
            double t = (double)d.tol;
            c2 = rr <= t * t * h->bb;
        } else {
            c2 = fbits(*slot) < d.tol;
        }
        // with residual replacement, only a replaced (true) residual stops the
        // solve: under the tolerance, or within twice the rounding floor of
        // an f32 iterate, (2^-24 / 2)^2 (|A| |x|)^2
        if (d.replace_every) c2 = replaced && (c2 || rr <= ATTAINABLE2 * aa);
        __hip_atomic_store(&h->converged, c2 ? 1u : 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // converged: the host's next checkpoint (first, first + chunk, ...)
        // after this iteration stops the solve
        uint32_t first = d.wstate[WS_FIRST], chunk = d.chunk ? d.chunk : 1u;
        if (c2 && first != 0u) {
            uint32_t next = cycle + 1u;
            uint32_t stop = next <= first ? first : first + (next - first + chunk - 1u) / chunk * chunk;
            __hip_atomic_store(&d.wstate[WS_STOP], stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    stepCycle(d, cycle);
}
// cgFinish (:2884-2893), and the words of the reference's fine level the
// packed solve kept aside: fluidCg's vectors, the rhs (-r), the drift, and,
// when an iteration ran, the rims the last post-relaxation read
// (prolongFine's p)
template <int S> __global__ void __launch_bounds__(64) k_cgFinish(WorldGrid g) {
    WORLD(g);
    if (d.store != S) return;
    if (!d.full) return;
    const int ran = d.fluid_cycles > 0u ? 1 : 0;
    if (bid == 0 && threadIdx.x == 0) {
        uint32_t stop = d.wstate[WS_STOP];
        if (stop < d.fluid_cycles) d.wstate[WS_LAST] = d.head->cycles;
    }
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (!inWorld(d, c)) continue;
        uint32_t i = cellIndex(d, c), pi = packIndex(d, c);
        // fluidCg as the iterations left it: p is the last iteration run's
        float x = d.cx[pi];
        *(float4 *)&CGV(d, i, 0) = make_float4(x, d.cr[pi], pBuf(d, ld(&d.head->cycles))[pi], d.cq[pi]);
        FLD(d, 2u * i, 0) = x;
        CNT(d, i, G_REGION) = 0u;
        if (d.cflag[pi] & CF_FLUID) {
            FLD(d, 2u * i, 1) = -d.cr[pi];
            FLD(d, 2u * i + 1u, 0) = fnLoad<S>(d, 0, pi).y;
            if (ran) {
                float2 p = fnLoad<S>(d, 1, pi);
                v2 e = prolonged<S>(d, 0u, c);
                CNT(d, i, G_RIM) = ubits(p.x + e.x);
                CNT(d, i, G_RIM + 1u) = ubits(p.y + e.y);
            }
        }
    }
}


// ============================================================ the bodies' passes
// bodyStart (:3065-3070): the step's push sums emptied
__global__ void __launch_bounds__(64) k_bodyStart(WorldGrid g) {
    WORLD(g);
    if (!d.full || !d.body_on) return;
    for (uint32_t k = bid * 64u + threadIdx.x; k < MAX_BODIES * BODY_PUSH_WORDS; k += nbid * 64u) d.pushes[k] = 0u;
}
// seepLoads (:3132-3143): the liquid's pressure on free body b's cell c where
// it rests against what isn't liquid or air
__device__ void seepLoads(const Dev &d, i2 c, uint32_t b) {
    for (int k = 0; k < 4; k++) {
        const i2 q = beside(c, k);
        if (!inWorld(d, q) || rigidAt(d, q) == b || gridZ(d, q) >= 0.0f) continue;
        const i2 n = I2(q.x - c.x, q.y - c.y);
        const v2 x = V2((float)c.x + 0.5f + 0.5f * (float)n.x, (float)c.y + 0.5f + 0.5f * (float)n.y);
        const v3 j = bodyFaceJ(d, q, c, b);
        const float s = seepAt(d, c, n, b, false, x);
        addPush(d, b, PUSH_SURFACE, V3(j.x * s, j.y * s, j.z * s));
    }
}
// bodyLoads (:3097-3121): the liquid's pressure where each free body's pose
// puts its surface, past what the cell's middle gives
__global__ void __launch_bounds__(64) k_bodyLoads(WorldGrid g) {
    WORLD(g);
    if (!d.full || !d.body_on) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        const i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (!inWorld(d, c)) continue;
        const uint32_t own = rigidAt(d, c);
        if (own != WNONE) {
            seepLoads(d, c, own);
            continue;
        }
        const float4 g4 = gridAt4(d, c);
        if (g4.z <= 0.0f) continue;
        const v2 mid = V2((float)c.x + 0.5f, (float)c.y + 0.5f);
        for (int k = 0; k < 4; k++) {
            const i2 q = beside(c, k);
            const uint32_t b = rigidAt(d, q);
            if (b == WNONE) continue;
            const i2 n = I2(q.x - c.x, q.y - c.y);
            const float reach = 1.0f - bodyReach(d, b, I2(-n.x, -n.y));
            const v2 surface = V2(mid.x + (float)n.x * reach, mid.y + (float)n.y * reach);
            const v3 j = bodyFaceJ(d, c, q, b);
            const float dp = potentialAt(d, mid) - potentialAt(d, surface);
            addPush(d, b, PUSH_SURFACE, V3(j.x * g4.z * dp, j.y * g4.z * dp, j.z * g4.z * dp));
        }
    }
}
// bodyDrag (:3150-3173): the liquid's drag on each free body as the
// viscosity's sweeps left its faces
__global__ void __launch_bounds__(64) k_bodyDrag(WorldGrid g) {
    WORLD(g);
    if (!d.full || !d.body_on || !d.visc_on) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        const i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        const uint32_t b = rigidAt(d, c);
        if (!inWorld(d, c) || b == WNONE) continue;
        const v2 pos = bodyV(d, b, BW_POS);
        const float rho = bodyDensity(d, b);
        v3 push = V3(0.0f, 0.0f, 0.0f);
        for (int k = 0; k < 4; k++) {
            const bool isU = k < 2;
            const int s = (k & 1) == 1 ? 1 : -1;
            const i2 n = isU ? I2(s, 0) : I2(0, s);
            if (gridZ(d, I2(c.x + n.x, c.y + n.y)) <= 0.0f) continue;
            const i2 at = isU ? I2(c.x + (k & 1), c.y) : I2(c.x, c.y + (k & 1));
            const v2 x = faceMiddle(at, isU);
            const v2 mv = bodyMoving(d, b, x, false);
            const float du = bodyReach(d, b, n) * rho * (FLD(d, faceSlot(d, faceIndex(d, at, isU)), 0) - (isU ? mv.x : mv.y));
            const v2 r = V2(x.x - pos.x, x.y - pos.y);
            if (isU) push.x += du;
            else push.y += du;
            push.z += isU ? -r.y * du : r.x * du;
        }
        if (push.x != 0.0f || push.y != 0.0f || push.z != 0.0f) addPush(d, b, PUSH_DRAG, push);
    }
}
// bodyFaces (:3180-3201): the bodies' faces, as the sweeps left them, moving
// with the bodies again, the drag and all
__global__ void __launch_bounds__(64) k_bodyFaces(WorldGrid g) {
    WORLD(g);
    if (!d.full || !d.body_on || !d.visc_on) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        const i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (!inWorld(d, c)) continue;
        Face3 fs[4];
        facesOf(d, c, fs);
        for (int j = 0; j < 4; j++) {
            if (fs[j].isU < 0) continue;
            const i2 at = I2(fs[j].x, fs[j].y);
            const bool isU = fs[j].isU == 1;
            const i2 a = isU ? I2(at.x - 1, at.y) : I2(at.x, at.y - 1);
            const uint32_t ra = rigidAt(d, a), rb = rigidAt(d, at);
            if (ra != WNONE || rb != WNONE) setBodyFace(d, faceIndex(d, at, isU), at, isU, gridAt4(d, a), gridAt4(d, at), ra, rb, true);
        }
    }
}
// cgBodyX, cgBodyP, cgBodyF (:2819-2845): the pressure's push on each free
// body: of the pressure the solve starts from, of an iteration's direction
// (into the iteration's set by its parity: cgDirection's p, newP, worked
// out again here as cgApply works it out), of the pressure it ends with
__global__ void __launch_bounds__(64) k_cgBodyX(WorldGrid g) {
    WORLD(g);
    if (!d.full || !d.body_on) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        const i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (inWorld(d, c) && gridZ(d, c) > 0.0f) addPressurePush(d, c, gridX(d, c), PUSH_X);
    }
}
template <int S> __global__ void __launch_bounds__(64) k_cgBodyP(WorldGrid g) {
    WORLD(g);
    if (d.store != S || !d.body_on) return;
    const Iter it = iterWords(d);
    const double beta = ld_f64(&d.head->beta);
    if (!it.encoded || it.converged) return;
    const float *pold = pBuf(d, it.cycle);
    for (uint32_t wid = bid; wid < it.active; wid += nbid) {
        const i2 c = blockCell(activeBlockAt(d, d.blocks[ACTIVE_SLOT(d, wid)]), threadIdx.x);
        if (!inWorld(d, c)) continue;
        const uint32_t pi = packIndex(d, c);
        if ((d.cflag[pi] & CF_BODY) == 0u) continue;
        addPressurePush(d, c, newP<S>(d, pi, beta, pold), PUSH_P0 + it.cycle % 2u);
    }
}
__global__ void __launch_bounds__(64) k_cgBodyF(WorldGrid g) {
    WORLD(g);
    if (!d.full || !d.body_on) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        const i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (inWorld(d, c) && gridZ(d, c) > 0.0f) addPressurePush(d, c, gridX(d, c), PUSH_F);
    }
}
// rigidify (:3037-3063): what the liquid's pressure did to each held body,
// and whether it is wet, into its liquid slot's words
__global__ void __launch_bounds__(64) k_rigidify(WorldGrid g) {
    WORLD(g);
    if (!d.full || !d.body_on || !d.bodies_run) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        const i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        const uint32_t h = heldAt(d, c);
        if (h == WNONE) continue;
        int32_t dd[3];
        heldPush(d, c, h, dd);
        if ((d.blocks[STATE_SLOT(d, blockIndex(d, idiv8(c)))] & NEAR_LIQUID) != 0u) atomicOr(&d.motion[liquidSlot(h, 3u)], LIQUID_WET);
        for (uint32_t k = 0; k < 3u; k++) atomicAdd(&d.motion[liquidSlot(h, k)], (uint32_t)dd[k]);
    }
}
// bodyPush (:3206-3218): what the liquid did to each free body this step
__global__ void __launch_bounds__(1024) k_bodyPush(WorldGrid g) {
    WORLD(g);
    if (!d.full || !d.body_on || !d.bodies_run) return;
    for (uint32_t b = threadIdx.x; b < MAX_BODIES; b += blockDim.x) {
        const v3 s = pushChangeFree(d, b, PUSH_SURFACE), dr = pushChangeFree(d, b, PUSH_DRAG), f = pushChangeFree(d, b, PUSH_F);
        const v3 ch = V3(s.x + dr.x + f.x, s.y + dr.y + f.y, s.z + dr.z + f.z);
        if (ch.x == 0.0f && ch.y == 0.0f && ch.z == 0.0f) continue;
        d.motion[liquidSlot(b, 0u)] += (uint32_t)wi32(wround(ch.x * RIGID_FIX));
        d.motion[liquidSlot(b, 1u)] += (uint32_t)wi32(wround(ch.y * RIGID_FIX));
        d.motion[liquidSlot(b, 2u)] += (uint32_t)wi32(wround(ch.z * RIGID_FIX));
    }
}

// ============================================================ after the solve
// project (:2895-2942)
__device__ void projectFace(const Dev &d, i2 at, bool isU) {
    uint64_t f = faceSlot(d, faceIndex(d, at, isU));
    i2 a = isU ? I2(at.x - 1, at.y) : I2(at.x, at.y - 1);
    float4 ga = gridAt4(d, a), gb = gridAt4(d, at);
    float beta = fineBeta(ga.z, gb.z, densityAt(d, a), densityAt(d, at));
    const uint32_t ra = rigidAt(d, a), rb = rigidAt(d, at);
    if (beta > 0.0f) {
        FLD(d, f, 0) = FLD(d, f, 0) - beta * (gb.x - ga.x);
        FLD(d, f, 2) = PROJECTED;
        // no drift into or out of a body's cells (:2907-2909)
        FLD(d, f, 3) = ra != WNONE || rb != WNONE ? 0.0f : wclamp(-beta * (driftAt(d, at) - driftAt(d, a)), -MAX_DRIFT, MAX_DRIFT);
    }
    // a face between the liquid and a free body moves as the body does with
    // the pressure's push the solve found (:2911-2921)
    if ((ra == WNONE) != (rb == WNONE) && (ra != WNONE ? gb.z : ga.z) > 0.0f) {
        const uint32_t b = ra != WNONE ? ra : rb;
        const v3 ch = pushChange(d, b, PUSH_F);
        const v2 m = faceMiddle(at, isU), pos = bodyV(d, b, BW_POS);
        const v2 r = V2(m.x - pos.x, m.y - pos.y);
        FLD(d, f, 0) += isU ? ch.x - ch.z * r.y : ch.y + ch.z * r.x;
        FLD(d, f, 2) = PROJECTED;
    }
}
__global__ void __launch_bounds__(64) k_project(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    if (bid == 0 && threadIdx.x == 0)
        atomicMax(&d.blocks[COUNT_SLOT(d, COUNT_MOST_CYCLES)], d.blocks[COUNT_SLOT(d, COUNT_CYCLES)]);
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (!inWorld(d, c)) continue;
        Face3 fs[4];
        facesOf(d, c, fs);
        for (int j = 0; j < 4; j++)
            if (fs[j].isU >= 0) projectFace(d, I2(fs[j].x, fs[j].y), fs[j].isU == 1);
    }
}
// extrapolate, extrapolateFurther (:3220-3275).  A face written here holds z
// under what its neighbours look for, so the pass reads the same in any
// order.
__global__ void __launch_bounds__(64) k_extrapolate(WorldGrid g, int layer) {
    WORLD(g);
    if (!d.full) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (!inWorld(d, c)) continue;
        Face3 fs[4];
        facesOf(d, c, fs);
        for (int j = 0; j < 4; j++) {
            if (fs[j].isU < 0) continue;
            bool isU = fs[j].isU == 1;
            i2 at = I2(fs[j].x, fs[j].y);
            uint64_t f = faceSlot(d, faceIndex(d, at, isU));
            i2 a = isU ? I2(at.x - 1, at.y) : I2(at.x, at.y - 1);
            if (FLD(d, f, 2) != 1.0f || gridZ(d, a) != 0.0f || gridZ(d, at) != 0.0f) continue;
            float reached = layer == 2 ? 1.5f : PROJECTED;
            float sum = 0.0f, n = 0.0f;
            for (int k = 0; k < 4; k++) {
                i2 q = beside(at, k);
                if (!faceInWorld(d, q, isU)) continue;
                uint64_t g = faceSlot(d, faceIndex(d, q, isU));
                float gz = FLD(d, g, 2);
                if (gz >= reached) {
                    sum += FLD(d, g, 0);
                    n += 1.0f;
                }
            }
            if (n > 0.0f) {
                FLD(d, f, 0) = sum / n;
                FLD(d, f, 2) = layer == 2 ? 1.25f : 1.5f;
            }
        }
    }
}

// advect (:3468-3694)
struct Sample { float now, before, weight, drift; v2 gradient; };
__device__ Sample sampleAxis(const Dev &d, v2 s, bool isU) {
    float fx0 = floorf(s.x), fy0 = floorf(s.y);
    i2 b = I2(wi32(fx0), wi32(fy0));
    v2 f = V2(s.x - fx0, s.y - fy0);
    float ax = 0.0f, ay = 0.0f;
    Sample r;
    r.now = r.before = r.weight = r.drift = 0.0f;
    int held = 0;
    v2 g = V2(0.0f, 0.0f);
    for (int j = 0; j < 4; j++) {
        int ox = j & 1, oy = j >> 1;
        float wgt = (ox == 1 ? f.x : 1.0f - f.x) * (oy == 1 ? f.y : 1.0f - f.y);
        v2 dw = V2((ox == 1 ? 1.0f : -1.0f) * (oy == 1 ? f.y : 1.0f - f.y), (ox == 1 ? f.x : 1.0f - f.x) * (oy == 1 ? 1.0f : -1.0f));
        i2 at = I2(b.x + ox, b.y + oy);
        if (!faceInWorld(d, at, isU)) continue;
        float4 fc = *(const float4 *)&FLD(d, faceSlot(d, faceIndex(d, at, isU)), 0);
        r.drift += wgt * fc.w;
        if (fc.z > 0.0f) {
            ax += wgt * fc.x;
            ay += wgt * fc.y;
            r.weight += wgt;
            g.x += dw.x * fc.x;
            g.y += dw.y * fc.x;
            held += fc.z == PROJECTED ? 1 : 0;
        }
    }
    if (r.weight > 0.0f) {
        r.now = ax / r.weight;
        r.before = ay / r.weight;
    }
    r.gradient = held == 4 ? g : V2(0.0f, 0.0f);
    return r;
}
__device__ __forceinline__ v2 clampSpeed(v2 v) {
    float s = wlength(v);
    return s > LIQUID_MAX_SPEED ? V2(v.x * (LIQUID_MAX_SPEED / s), v.y * (LIQUID_MAX_SPEED / s)) : v;
}
__device__ __forceinline__ bool escapesTo(const Dev &d, i2 q, uint32_t side) {   // :3535-3541
    if (!inWorld(d, q) || shutAt(d, q)) return false;
    const uint32_t cup = cupOf(d, cellIndex(d, q));
    return cup == WNONE || cup == side;
}
__device__ v2 escape(const Dev &d, v2 pos) {                                 // :3552-3594
    if (!isObstaclePos(d, pos)) return pos;
    i2 c = I2(wi32(floorf(pos.x)), wi32(floorf(pos.y)));
    const uint32_t side = inWorld(d, c) ? cupBefore(d, cellIndex(d, c)) : WNONE;
    v2 best = pos;
    float bestD = 1e9f;
    for (int n = 0; n < 8; n++) {
        int m = n + (n >= 4 ? 1 : 0);
        i2 at = I2(c.x + m % 3 - 1, c.y + m / 3 - 1);
        if (!escapesTo(d, at, side)) continue;
        v2 inside = V2(wclamp(pos.x, (float)at.x + 0.02f, (float)at.x + 0.98f), wclamp(pos.y, (float)at.y + 0.02f, (float)at.y + 0.98f));
        float dd = wlength(V2(pos.x - inside.x, pos.y - inside.y));
        if (dd < bestD) { bestD = dd; best = inside; }
    }
    if (bestD < 1e9f) return best;
    int32_t kt = downTurns(d, c);
    i2 ways[4] = {I2(0, -1), I2(-1, 0), I2(1, 0), I2(0, 1)};
    i2 way = turnStep(ways[0], kt);
    int nearest = ESCAPE_REACH + 1;
    for (int n = 0; n < 4; n++) {
        i2 dir = turnStep(ways[n], kt);
        for (int r = 2; r < nearest; r++) {
            i2 q = I2(c.x + dir.x * r, c.y + dir.y * r);
            if (!isObstaclePos(d, V2((float)q.x + 0.5f, (float)q.y + 0.5f)) && escapesTo(d, q, side)) {
                nearest = r;
                way = dir;
                break;
            }
        }
    }
    return inWorld(d, I2(c.x + way.x, c.y + way.y)) ? V2(pos.x + (float)way.x, pos.y + (float)way.y) : pos;
}
__global__ void __launch_bounds__(256) k_advect(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    uint32_t cnt = ld(&d.parts[0]);
    uint32_t n = cnt < d.cap ? cnt : d.cap;
    const uint32_t lane = lane_id();
    for (uint32_t base = bid * blockDim.x + (threadIdx.x & ~63u); base < n; base += nbid * blockDim.x) {   // (this wavefront's)
        const uint32_t k = base + lane;
        uint32_t came = 0xffffffffu, how = 0u;   // the cell arrived in and its came bit
        if (k < n) {
        uint4 mo = *(const uint4 *)&MOTION(d, d.parts, k, 0);
        uint4 st = *(const uint4 *)&STATE(d, d.parts, k, 0);
        v2 orig = V2(fbits(mo.x), fbits(mo.y));
        v2 qvel = V2(fbits(mo.z), fbits(mo.w));
        float temp = fbits(st.y);
        v2 qpos = escape(d, orig);
        Sample u = sampleAxis(d, V2(qpos.x - 0.0f, qpos.y - 0.5f), true);
        Sample v = sampleAxis(d, V2(qpos.x - 0.5f, qpos.y - 0.0f), false);
        v2 flow = V2(u.weight > 0.0f ? u.now : qvel.x, v.weight > 0.0f ? v.now : qvel.y);
        v2 vel = clampSpeed(flow);
        float4 aff = make_float4(AFFINE_KEEP * u.gradient.x, AFFINE_KEEP * u.gradient.y, AFFINE_KEEP * v.gradient.x, AFFINE_KEEP * v.gradient.y);
        // slipOf (:3605-3614)
        v2 slip = V2(0.0f, 0.0f);
        {
            uint32_t i = cellIndex(d, cellOf(d, qpos));
            uint32_t nn = CNT(d, i, G_COUNT);
            if (nn >= 2u) {
                float mean = (float)CNT(d, i, G_MASS) / (FIXF * (float)nn);
                float lighter = (mean - liquidDensity(temp)) / fmaxf(mean, 1e-3f);
                const v2 pull = pullAt(d, qpos);
                slip = V2(-pull.x / WATER_WEIGHT * SLIP * lighter, -pull.y / WATER_WEIGHT * SLIP * lighter);
            }
        }
        v2 cs = clampSpeed(vel);
        v2 travel = V2(cs.x + wclamp(u.drift + slip.x, -MAX_DRIFT, MAX_DRIFT), cs.y + wclamp(v.drift + slip.y, -MAX_DRIFT, MAX_DRIFT));
        int32_t steps = wi32(ceilf(fmaxf(fabsf(travel.x), fabsf(travel.y)) * 2.0f));
        steps = steps < 1 ? 1 : (steps > ADVECT_STEPS ? ADVECT_STEPS : steps);
        v2 dd = V2(travel.x / (float)steps, travel.y / (float)steps);
        v2 pos = qpos;
        for (int32_t s = 0; s < steps; s++) {
            v2 nx = V2(pos.x + dd.x, pos.y + 0.0f);
            if (dd.x != 0.0f && isObstaclePos(d, nx)) {
                if (vel.x * dd.x > 0.0f) vel.x = 0.0f;
                dd.x = 0.0f;
            } else {
                pos = nx;
            }
            v2 ny = V2(pos.x + 0.0f, pos.y + dd.y);
            if (dd.y != 0.0f && isObstaclePos(d, ny)) {
                if (vel.y * dd.y > 0.0f) vel.y = 0.0f;
                dd.y = 0.0f;
            } else {
                pos = ny;
            }
        }
        if (!isActiveBlock(d, idiv8(cellOf(d, pos)))) {
            pos = qpos;
            if (!isActiveBlock(d, idiv8(cellOf(d, pos)))) pos = orig;
        }
        *(uint4 *)&MOTION(d, d.parts, k, 0) = make_uint4(ubits(pos.x), ubits(pos.y), ubits(vel.x), ubits(vel.y));
        *(float4 *)&AFFINE(d, d.parts, k, 0) = aff;
        uint32_t li = 2u * cellIndex(d, cellOf(d, qpos));
        float lx = FLD(d, li, 0), lz = FLD(d, li, 2);
        if (lz > 0.0f) {
            bool falling = lx < 0.5f * lz * wlength(pullAt(d, qpos));
            came = cellIndex(d, cellOf(d, pos));
            how = falling ? CAME_FALLING : CAME_PRESSING;
        }
        }
        // the came bits OR-ed over each run of lanes arriving in one cell,
        // one atomic a run (OR is indifferent to order and repeats)
        const uint32_t before = __shfl_up(came, 1, WAVE);
        const bool head = lane == 0u || before != came;
        const uint64_t heads = __ballot(head);
        const uint64_t after = heads & (lane == 63u ? 0ull : ~((1ull << (lane + 1u)) - 1ull));
        const uint32_t next = after ? (uint32_t)__ffsll((long long)after) - 1u : 64u;
        for (uint32_t off = 1; off < 64u; off <<= 1) {
            const uint32_t o = __shfl_down(how, off, WAVE);
            if (lane + off < next) how |= o;
        }
        if (head && came != 0xffffffffu && how) atomicOr(&CNT(d, came, G_CAME), how);
    }
}

// ================================================================ sort
// sortCount (:3700-3708)
// (each particle's place among its cell's: a run of lanes of one cell in a
// wavefront takes its places with one atomic, the particles being mostly in
// cell order already; the order within a cell is the atomics' either way)
__global__ void __launch_bounds__(256) k_sortCount(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    uint32_t cnt = ld(&d.parts[0]);
    uint32_t n = cnt < d.cap ? cnt : d.cap;
    const uint32_t lane = lane_id();
    const uint64_t upTo = lane == 63u ? ~0ull : ((1ull << (lane + 1u)) - 1ull);   // lanes 0 .. this one
    for (uint32_t base = bid * blockDim.x + (threadIdx.x & ~63u); base < n; base += nbid * blockDim.x) {   // (this wavefront's)
        const uint32_t k = base + lane;
        const bool on = k < n;
        uint32_t cell = 0xffffffffu;
        if (on) cell = cellIndex(d, cellOf(d, V2(fbits(MOTION(d, d.parts, k, 0)), fbits(MOTION(d, d.parts, k, 1)))));
        const uint32_t before = __shfl_up(cell, 1, WAVE);
        const bool head = lane == 0u || before != cell;
        const uint64_t heads = __ballot(head);
        const uint32_t start = 63u - (uint32_t)__clzll((long long)(heads & upTo));
        const uint64_t after = heads & ~upTo;
        const uint32_t next = after ? (uint32_t)__ffsll((long long)after) - 1u : 64u;
        uint32_t first = 0;
        if (head && on) first = atomicAdd(&CNT(d, cell, G_SORT), next - lane);
        first = __shfl(first, (int)start, WAVE);
        if (on) STATE(d, d.parts, k, 3) = first + (lane - start);
    }
}
// scanBlocks (:3730-3753)
__global__ void __launch_bounds__(64) k_scanBlocks(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        uint32_t n = inWorld(d, c) ? CNT(d, cellIndex(d, c), G_SORT) : 0u;
        uint32_t incl = wave_incl_scan(n);
        if (inWorld(d, c)) {
            uint32_t i = cellIndex(d, c);
            CNT(d, i, G_START) = incl - n;
            uint32_t came = CNT(d, i, G_CAME);
            CNT(d, i, G_CAME) = 0u;
            if (came == CAME_FALLING) FLD(d, 2u * i + 1u, 3) = (float)(wu32(FLD(d, 2u * i + 1u, 3)) | fallingBit(d, 0u));
        }
        if (threadIdx.x == 63) d.blocks[TOTAL_SLOT(d, wid)] = incl;
    }
}
// A single-workgroup scan's thread takes a contiguous range of entries; the
// first RANGE_HELD of them are loaded at once into registers (the loads
// below the range's end, made together) and kept for the second pass;
// entries past those are read one by one.
#define RANGE_HELD 32
// scanTotals (:3757-3780): the blocks' runs' starts, in the active list's order
__global__ void __launch_bounds__(256) k_scanTotals(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    __shared__ uint32_t s[256];
    uint32_t li = threadIdx.x, n = activeCount(d);
    uint32_t per = (n + 255u) / 256u, lo = li * per < n ? li * per : n, hi = lo + per < n ? lo + per : n;
    uint32_t held[RANGE_HELD];
#pragma unroll
    for (int q = 0; q < RANGE_HELD; q++) held[q] = lo + (uint32_t)q < hi ? d.blocks[TOTAL_SLOT(d, lo + (uint32_t)q)] : 0u;
    uint32_t sum = 0;
#pragma unroll
    for (int q = 0; q < RANGE_HELD; q++) sum += held[q];
    for (uint32_t j = lo + RANGE_HELD; j < hi; j++) sum += d.blocks[TOTAL_SLOT(d, j)];
    s[li] = sum;
    __syncthreads();
    for (uint32_t o = 1; o < 256u; o <<= 1) {
        uint32_t a = li >= o ? s[li - o] : 0u;
        __syncthreads();
        s[li] += a;
        __syncthreads();
    }
    uint32_t run = s[li] - sum;
#pragma unroll
    for (int q = 0; q < RANGE_HELD; q++) {
        if (lo + (uint32_t)q >= hi) break;
        d.blocks[TOTAL_SLOT(d, lo + (uint32_t)q)] = run;
        run += held[q];
    }
    for (uint32_t j = lo + RANGE_HELD; j < hi; j++) {
        uint32_t t = d.blocks[TOTAL_SLOT(d, j)];
        d.blocks[TOTAL_SLOT(d, j)] = run;
        run += t;
    }
    if (li == 255u) {
        d.sorted[0] = s[255];
        d.blocks[COUNT_SLOT(d, COUNT_PARTICLES)] = s[255];
        // the run totals (this library's): once a full step
        uint32_t pc = d.parts[0];
        d.totals[0] += 1ull;
        d.totals[1] += (unsigned long long)d.blocks[COUNT_SLOT(d, COUNT_CYCLES)];
        d.totals[2] += (unsigned long long)(pc < d.cap ? pc : d.cap);
        d.totals[3] += 64ull * n;
    }
}
// scanFinish (:3783-3790)
__global__ void __launch_bounds__(64) k_scanFinish(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (inWorld(d, c)) CNT(d, cellIndex(d, c), G_START) += d.blocks[TOTAL_SLOT(d, wid)];
    }
}
// scatter (:3793-3808)
__global__ void __launch_bounds__(256) k_scatter(WorldGrid g) {
    WORLD(g);
    if (!d.full) return;
    uint32_t cnt = ld(&d.parts[0]);
    uint32_t n = cnt < d.cap ? cnt : d.cap;
    for (uint32_t k = bid * blockDim.x + threadIdx.x; k < n; k += nbid * blockDim.x) {
        uint4 mo = *(const uint4 *)&MOTION(d, d.parts, k, 0);
        uint4 st = *(const uint4 *)&STATE(d, d.parts, k, 0);
        uint4 af = *(const uint4 *)&AFFINE(d, d.parts, k, 0);
        uint32_t i = cellIndex(d, cellOf(d, V2(fbits(mo.x), fbits(mo.y))));
        uint32_t m = CNT(d, i, G_START) + st.w;
        if (m < d.cap) {
            *(uint4 *)&MOTION(d, d.sorted, m, 0) = mo;
            *(uint4 *)&STATE(d, d.sorted, m, 0) = st;
            *(uint4 *)&AFFINE(d, d.sorted, m, 0) = af;
        }
    }
}

// ================================================================ draw
__device__ __forceinline__ bool vacated(const Dev &d, uint32_t j) {
    return kindOf(d, wType(d.cells[4ull * j])) == KIND_LIQUID && CNT(d, j, G_SORT) == 0u;
}
__device__ uint32_t airTo(const Dev &d, i2 c, bool gas) {                    // :3898-3924
    if (!gas && gridZ(d, c) > 0.0f) return WNONE;
    int32_t kt = downTurns(d, c);
    for (int n = 0; n < 8; n++) {
        i2 o;
        if (!gas) {
            const int ax[8] = {0, -1, 1, 0, -1, 1, -1, 1}, ay[8] = {-1, 0, 0, 1, -1, -1, 1, 1};
            o = I2(ax[n], ay[n]);
        } else {
            const int gx[8] = {0, -1, 1, -1, 1, -1, 1, 0}, gy[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
            o = I2(gx[n], gy[n]);
        }
        o = turnStep(o, kt);
        i2 at = I2(c.x + o.x, c.y + o.y);
        if (!inWorld(d, at)) continue;
        uint32_t j = cellIndex(d, at);
        if (vacated(d, j) || (gas && wType(d.cells[4ull * j]) == E_AIR && CNT(d, j, G_SORT) == 0u && !isOccupant(occupantAt(d, j)))) return j;
    }
    return WNONE;
}
__global__ void __launch_bounds__(64) k_rasterClaims(WorldGrid g) {               // :3928-3948
    WORLD(g);
    if (!d.full) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (!inWorld(d, c)) continue;
        uint32_t i = cellIndex(d, c), to = WNONE;
        uint32_t t = wType(d.cells[4ull * i]);
        if (CNT(d, i, G_SORT) > 0u && !isOccupant(occupantAt(d, i))) {
            if (t == E_AIR) to = airTo(d, c, false);
            else if (kindOf(d, t) == KIND_GAS && gridZ(d, c) >= 0.0f) to = airTo(d, c, true);
        }
        CNT(d, i, G_AIR_TO) = to;
        if (to != WNONE) atomicMax(&CNT(d, to, G_TAKEN), ~i);
    }
}
__device__ __forceinline__ void setGasPair(const Dev &d, uint32_t i, uint32_t a, uint32_t b) {
    uint64_t at = gasWord(d, i);
    d.cells[at] = a;
    d.cells[at + 1] = b;
}
// pushGasOut (:3855-3861; airsource.wgsl:65-72): cell i's gas into its
// airSource slot, with gasHeatCap (common.wgsl:473-475) times temp, and none
// left in the cell
__device__ void pushGasOut(const Dev &d, uint32_t i, float temp) {
    const uint64_t at = gasWord(d, i);
    const v2 m = unpack2x16(d.cells[at]), r = unpack2x16(d.cells[at + 1]);
    const float vapour = m.x, mist = m.y, lost = r.x, amount = r.y;
    addSource(d, i, SOURCE_AMOUNT, amount);
    addSource(d, i, SOURCE_VAPOUR, vapour);
    addSource(d, i, SOURCE_MIST, mist);
    addSource(d, i, SOURCE_LOST, lost * amount);
    addSource(d, i, SOURCE_HEAT, (AIR_HEATCAP * amount + (VAPOUR_HEAT_CAP * vapour + WATER_HEATCAP * mist) / WATER_VAPOUR) * temp);
    setGasPair(d, i, 0u, 0u);
}
__global__ void __launch_bounds__(64) k_raster(WorldGrid g) {                      // :3983-4119
    WORLD(g);
    if (!d.full) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (!inWorld(d, c)) continue;
        uint32_t i = cellIndex(d, c), n = CNT(d, i, G_SORT);
        if (n == 0u) continue;
        uint32_t start = CNT(d, i, G_START);
        uint4 v = *(const uint4 *)&d.cells[4ull * i];
        uint32_t t = wType(v.x);
        // a person's body: not drawn, its particles left to find their way
        // out, and liquid drawn there before the person came turned air
        if (isOccupant(occupantAt(d, i))) {
            if (kindOf(d, t) == KIND_LIQUID) {
                *(uint4 *)&d.cells[4ull * i] = make_uint4(E_AIR, ubits(effTemp(fbits(v.y))), 0u, 0u);
                setGasPair(d, i, 0u, 0u);
            }
            CNT(d, i, G_DEST) = DEST_KEEP;
            continue;
        }
        bool drawn = false;
        if (kindOf(d, t) == KIND_LIQUID) {
            drawn = true;
        } else if (t == E_AIR || (kindOf(d, t) == KIND_GAS && gridZ(d, c) >= 0.0f)) {
            // moveOutOfTheWay (:3954-3973)
            uint32_t to = CNT(d, i, G_AIR_TO);
            if (to != WNONE && CNT(d, to, G_TAKEN) == ~i) {
                // pushGasOut at the air's temperature as stored (AIR has no
                // phase points: js/elements.js:424-426), or, without the air, let go
                if (wType(d.cells[4ull * to]) == E_AIR) {
                    if (d.air_on) pushGasOut(d, to, fbits(d.cells[4ull * to + 1]));
                    else setGasPair(d, to, 0u, 0u);
                }
                *(uint4 *)&d.cells[4ull * to] = v;
                uint64_t gi = gasWord(d, i);
                setGasPair(d, to, d.cells[gi], d.cells[gi + 1]);
                setGasPair(d, i, 0u, 0u);
                drawn = true;
            } else if (t == E_AIR) {
                if (d.air_on) pushGasOut(d, i, fbits(v.y));
                else setGasPair(d, i, 0u, 0u);
                drawn = true;
            }
        }
        if (!drawn) {
            CNT(d, i, G_DEST) = DEST_KEEP;
            continue;
        }
        uint32_t kinds[RASTER_KINDS] = {WNONE, WNONE, WNONE, WNONE}, counts[RASTER_KINDS] = {0, 0, 0, 0};
        for (uint32_t k = start; k < start + n; k++) {
            uint32_t pt = wType(STATE(d, d.sorted, k, 0)), at = RASTER_KINDS;
            for (uint32_t m = 0; m < RASTER_KINDS; m++)
                if (kinds[m] == pt) at = m;
            if (at == RASTER_KINDS && pt < kinds[RASTER_KINDS - 1u]) {
                at = RASTER_KINDS - 1u;
                kinds[at] = pt;
                counts[at] = 0u;
                while (at > 0u && kinds[at - 1u] > kinds[at]) {
                    uint32_t t0 = kinds[at - 1u], c0 = counts[at - 1u];
                    kinds[at - 1u] = kinds[at];
                    counts[at - 1u] = counts[at];
                    kinds[at] = t0;
                    counts[at] = c0;
                    at--;
                }
            }
            if (at < RASTER_KINDS) counts[at]++;
        }
        uint32_t most = kinds[0], votes = counts[0];
        for (uint32_t m = 1; m < RASTER_KINDS; m++)
            if (counts[m] > votes) { most = kinds[m]; votes = counts[m]; }
        uint32_t word = 0u, lowest = WNONE, solute = 0u, soluteId = WNONE;
        bool solutes = most == E_WATER;
        for (uint32_t k = start; k < start + n; k++) {
            uint32_t sx = STATE(d, d.sorted, k, 0), sz = STATE(d, d.sorted, k, 2);
            if (wType(sx) == most && (lowest == WNONE || sz < lowest)) { word = sx; lowest = sz; }
            if (solutes && wType(sx) == most && wAux(sx) != 0u && (soluteId == WNONE || sz < soluteId)) { solute = wAux(sx); soluteId = sz; }
        }
        if (solutes) word = packWord(wType(word), solute, wLife(word), wShade(word), wFlags(word));
        float livesx = 0.0f, livesy = 0.0f;
        u2 heat = {0, 0}, capOf = {0, 0}, velX = {0, 0}, velY = {0, 0};
        for (uint32_t k = start; k < start + n; k++) {
            uint4 st = *(const uint4 *)&STATE(d, d.sorted, k, 0);
            uint32_t pt = wType(st.x);
            if (pt == most && wAux(st.x) == wAux(word)) { livesx += (float)wLife(st.x); livesy += 1.0f; }
            else if (pt == most && solutes && wAux(st.x) == 0u) { livesy += 1.0f; }
            float hc = fmaxf(WATER_HEATCAP, 1e-6f);
            heat = addExact(heat, hc * fbits(st.y));
            capOf = addExact(capOf, hc);
            velX = addExact(velX, fbits(MOTION(d, d.sorted, k, 2)));
            velY = addExact(velY, fbits(MOTION(d, d.sorted, k, 3)));
        }
        Rng r;
        seedRng(r, i, d.frame, d.seed ^ 0x7a57u);
        uint32_t life = wu32(floorf(livesx / livesy + rnd(r)));
        life = life < 255u ? life : 255u;
        word = packWord(wType(word), wAux(word), life, wShade(word), wFlags(word));
        float temp = exactValue(heat, EXACT_FINE) / exactValue(capOf, EXACT_FINE);
        v2 vel = V2(exactValue(velX, EXACT_FINE) / (float)n, exactValue(velY, EXACT_FINE) / (float)n);
        *(uint4 *)&d.cells[4ull * i] = make_uint4(wWith(word, FL_DRAWN), ubits(temp), pack2x16(vel), (i << 8) | (n < LIQUID_COUNT_MAX ? n : LIQUID_COUNT_MAX));
        CNT(d, i, G_TYPE) = (wType(word) | (wAux(word) << 8)) | (life << 16);
    }
}
__global__ void __launch_bounds__(64) k_unpaint(WorldGrid g) {                     // :4163-4182
    WORLD(g);
    if (!d.full) return;
    FOR_ACTIVE_BLOCKS(d, wid) {
        i2 c = blockCell(activeBlockAt(d, blk_), threadIdx.x);
        if (!inWorld(d, c)) continue;
        uint32_t i = cellIndex(d, c);
        uint4 v = *(const uint4 *)&d.cells[4ull * i];
        if (kindOf(d, wType(v.x)) == KIND_LIQUID && CNT(d, i, G_SORT) == 0u && CNT(d, i, G_TAKEN) == 0u) {
            *(uint4 *)&d.cells[4ull * i] = make_uint4(E_AIR, ubits(effTemp(fbits(v.y))), 0u, 0u);
            setGasPair(d, i, 0u, 0u);
        }
    }
}

// ============================================================ batches
// A convergence checkpoint of a batch: each world's stop (the iteration its
// solve stops before, NO_STOP while it runs) and the iterations run.
__global__ void __launch_bounds__(64) k_stops(WorldGrid g, uint32_t *out) {
    WORLD(g);
    if (threadIdx.x == 0) {
        out[2u * wi_] = ld(&d.wstate[WS_STOP]);
        out[2u * wi_ + 1u] = ld(&d.head->cycles);
    }
}
// A batch's reports (js/sim.js:1760-1780): each world's FLUID_STATS, then
// its MOST_CYCLES cleared as a report on steps does.
__global__ void __launch_bounds__(64) k_reports(WorldGrid g, uint32_t *out) {
    WORLD(g);
    uint32_t t = threadIdx.x;
    if (t < (uint32_t)FLUID_STATS) out[wi_ * (uint32_t)FLUID_STATS + t] = d.blocks[COUNT_SLOT(d, t)];
    __syncthreads();
    if (t == 0) d.blocks[COUNT_SLOT(d, COUNT_MOST_CYCLES)] = 0u;
}

// the kernels of the V-cycle's storage, one of each STORE_*
#define STORE_KERNELS(S) \
    template __global__ void k_relax<S>(WorldGrid, int); \
    template __global__ void k_restrict<S>(WorldGrid); \
    template __global__ void k_levelRho<S>(WorldGrid); \
    template __global__ void k_levelDown<S>(WorldGrid, uint32_t); \
    template __global__ void k_levelUp<S>(WorldGrid, uint32_t); \
    template __global__ void k_coarseLDS<S>(WorldGrid); \
    template __global__ void k_cgPrepare<S>(WorldGrid); \
    template __global__ void k_cgRho<S>(WorldGrid); \
    template __global__ void k_cgApply<S>(WorldGrid); \
    template __global__ void k_cgUpdate<S>(WorldGrid); \
    template __global__ void k_cgTrue<S>(WorldGrid); \
    template __global__ void k_cgFinish<S>(WorldGrid); \
    template __global__ void k_cgBodyP<S>(WorldGrid);
STORE_KERNELS(0)
STORE_KERNELS(1)
STORE_KERNELS(2)

// ============================================================ the FP64 check
// Plain CG in FP64 on the matrix of the world's last solve (its packed
// betas and pins, applied as cgApply applies them) for the rhs that solve
// captured, from 0: the reference a solve's pressure is measured against.
// Sums are per block (a wavefront's halving), then one workgroup's over the
// blocks in list order.  head64: rr, bb, pq, alpha, beta, then the check's
// four sums.
__device__ __forceinline__ double apply64(const Dev &d, i2 c, uint32_t i, const double *p) {
    float4 cf = d.coef[i];
    const float v[4] = {cf.x, cf.y, cf.z, cf.w};
    double sum = 0.0, diag = 0.0;
    for (int k = 0; k < 4; k++) {
        double b = (double)fabsf(v[k]);
        if (b == 0.0) continue;
        if (__float_as_uint(v[k]) >> 31) sum += b * p[packIndex(d, beside(c, k))];
        diag += b;
    }
    if (d.cflag[i] & CF_PINNED) diag += (double)PINF;
    return diag * p[i] - sum;
}
#define FLUID_CELL(d, c) (inWorld(d, c) && (d.cflag[packIndex(d, c)] & CF_FLUID))
__device__ void sum64(const Dev &d, int nv, double *out) {
    __shared__ double t[4][SUM_LANES];
    uint32_t li = threadIdx.x, na = activeCount(d);
    for (int k = 0; k < nv; k++) {
        double a = 0.0;
        for (uint32_t j = li; j < na; j += SUM_LANES) a += d.part64[4ull * j + k];
        t[k][li] = a;
    }
    __syncthreads();
    for (uint32_t h = SUM_LANES / 2u; h > 0; h >>= 1) {
        if (li < h)
            for (int k = 0; k < nv; k++) t[k][li] += t[k][li + h];
        __syncthreads();
    }
    if (li == 0)
        for (int k = 0; k < nv; k++) out[k] = t[k][0];
}
__global__ void __launch_bounds__(256) k64_start(WorldGrid g) {
    WORLD(g);
    FOR_ACTIVE_BLOCKS4(d, wid) {
        i2 c = blockCell(activeBlock(d, wid), LANE);
        double bb = 0.0;
        if (inWorld(d, c)) {
            uint32_t i = packIndex(d, c);
            double b = FLUID_CELL(d, c) ? (double)d.capb[i] : 0.0;
            d.x64[i] = 0.0;
            d.r64[i] = b;
            d.p64[i] = b;
            bb = b * b;
        }
        bb = wave_tree_sum(bb);
        if (LANE == 0) d.part64[4ull * wid] = bb;
    }
}
// which 0: rr = bb (the start); 1: pq, alpha; 2: the new rr, beta
__global__ void __launch_bounds__(1024) k64_sum(WorldGrid g, int which) {
    WORLD(g);
    double v[1];
    sum64(d, 1, v);
    if (threadIdx.x == 0) {
        double *h = d.head64;
        if (which == 0) {
            h[0] = v[0];
            h[1] = v[0];
        } else if (which == 1) {
            h[2] = v[0];
            h[3] = v[0] > 0.0 ? h[0] / v[0] : 0.0;
        } else {
            h[4] = h[0] > 0.0 ? v[0] / h[0] : 0.0;
            h[0] = v[0];
        }
    }
}
__global__ void __launch_bounds__(256) k64_apply(WorldGrid g) {
    WORLD(g);
    FOR_ACTIVE_BLOCKS4(d, wid) {
        i2 c = blockCell(activeBlock(d, wid), LANE);
        double pq = 0.0;
        if (FLUID_CELL(d, c)) {
            uint32_t i = packIndex(d, c);
            double q = apply64(d, c, i, d.p64);
            d.q64[i] = q;
            pq = d.p64[i] * q;
        }
        pq = wave_tree_sum(pq);
        if (LANE == 0) d.part64[4ull * wid] = pq;
    }
}
__global__ void __launch_bounds__(256) k64_update(WorldGrid g) {
    WORLD(g);
    double alpha = d.head64[3];
    FOR_ACTIVE_BLOCKS4(d, wid) {
        i2 c = blockCell(activeBlock(d, wid), LANE);
        double rr = 0.0;
        if (FLUID_CELL(d, c)) {
            uint32_t i = packIndex(d, c);
            d.x64[i] += alpha * d.p64[i];
            double r = d.r64[i] - alpha * d.q64[i];
            d.r64[i] = r;
            rr = r * r;
        }
        rr = wave_tree_sum(rr);
        if (LANE == 0) d.part64[4ull * wid] = rr;
    }
}
__global__ void __launch_bounds__(256) k64_direction(WorldGrid g) {
    WORLD(g);
    double beta = d.head64[4];
    FOR_ACTIVE_BLOCKS4(d, wid) {
        i2 c = blockCell(activeBlock(d, wid), LANE);
        if (FLUID_CELL(d, c)) {
            uint32_t i = packIndex(d, c);
            d.p64[i] = d.r64[i] + beta * d.p64[i];
        }
    }
}
// the check: |x - x64|^2, |x64|^2 for the world's float x (the solve's),
// then |b - A x|^2 and |b - A x64|^2 (p64 and q64 take x and x64 as float
// x in double)
__global__ void __launch_bounds__(256) k64_check(WorldGrid g, int phase) {
    WORLD(g);
    FOR_ACTIVE_BLOCKS4(d, wid) {
        i2 c = blockCell(activeBlock(d, wid), LANE);
        double a = 0.0, b2 = 0.0;
        if (inWorld(d, c)) {
            uint32_t i = packIndex(d, c);
            if (phase == 0) {
                d.p64[i] = FLUID_CELL(d, c) ? (double)d.cx[i] : 0.0;
            } else if (FLUID_CELL(d, c)) {
                double e = d.p64[i] - d.x64[i];
                a = e * e;
                b2 = d.x64[i] * d.x64[i];
            }
        }
        if (phase == 0) continue;
        a = wave_tree_sum(a);
        b2 = wave_tree_sum(b2);
        if (LANE == 0) {
            d.part64[4ull * wid] = a;
            d.part64[4ull * wid + 1] = b2;
        }
    }
}
__global__ void __launch_bounds__(256) k64_residual(WorldGrid g, int of_x64) {
    WORLD(g);
    const double *x = of_x64 ? d.x64 : d.p64;
    FOR_ACTIVE_BLOCKS4(d, wid) {
        i2 c = blockCell(activeBlock(d, wid), LANE);
        double rr = 0.0;
        if (FLUID_CELL(d, c)) {
            uint32_t i = packIndex(d, c);
            double r = (double)d.capb[i] - apply64(d, c, i, x);
            rr = r * r;
        }
        rr = wave_tree_sum(rr);
        if (LANE == 0) d.part64[4ull * wid] = rr;
    }
}
__global__ void __launch_bounds__(1024) k64_keep(WorldGrid g, int nv, int at) {
    WORLD(g);
    double v[4];
    sum64(d, nv, v);
    if (threadIdx.x == 0)
        for (int k = 0; k < nv; k++) d.head64[at + k] = v[k];
}
