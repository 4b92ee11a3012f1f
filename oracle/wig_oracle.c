/*
 * ORACLE — test infrastructure only.  Not part of the product: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * A CPU restatement, in plain C, of worldinagrain's liquid step for a world
 * of water, air and obstacles: shaders/fluid.wgsl as encoded by
 * Simulation.encodeFluid (js/sim.js:2466-2612), kernel by kernel in the
 * reference's order, each kernel a loop over its dispatch domain.  Citations
 * are to shaders/fluid.wgsl unless another file is named.
 *
 * The inputs the reference's other phases supply come in as parts, each
 * read only when its WigStepCtl switch is on (else what the reference
 * computes with none): the air (air[0], airSource), people's bodies
 * (occupants), the gravity pass's field, and the free and held bodies (the
 * bodies' table in `motion`, the swept words of the bodies buffer, and the
 * step's push sums).  Liquid arithmetic takes water's constants
 * (js/elements.js:449-457), fresh (no solute); viscosity and surface
 * tension by the element columns of wig_set_element_props.
 * Everything order-dependent in the reference is made order-free there
 * (fixed-point sums, exact sums, claims by atomicMax); where the reference's
 * order is arbitrary (the compaction of `parts`, the active list, ranks in a
 * cell's run) this restatement takes slot order and block-index order.
 *
 * Floating point follows WGSL: f32 throughout, round() ties to even, no
 * contraction (built with -ffp-contract=off).  The CG sums run in one of two
 * orders: the reference's f32 tree (dot_mode 0: fluid.wgsl:2628-2661) or
 * FP64 with the same tree shape (dot_mode 1), which is what the HIP path
 * computes.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/wig.h"

/* Planted mistakes (test infrastructure): built with -DORACLE_MUTANT=k
 * (oracle/Makefile's `mutants`), this restatement carries mistake k, one a
 * build, which the pins in tests/ must catch (tests/test_oracle_mutants.py).
 * 0, the default, plants none. */
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif
#define MUTANT(k) (ORACLE_MUTANT == (k))

/* Built with -fopenmp (oracle/Makefile's `omp`, the CPU baseline's build),
 * the loops the reference dispatches as independent invocations run on all
 * cores; where two invocations meet in one word the reference uses an
 * atomic, and so does this build (at_*), on integers, so every result is
 * the serial build's (tests/test_oracle.py checks).  Union-find, the block
 * lists and the fixed-order sums stay serial. */
#ifdef _OPENMP
#define PAR_FOR _Pragma("omp parallel for schedule(static)")
#define PAR_FOR_MOST _Pragma("omp parallel for schedule(static) reduction(max: most)")
static inline void at_add(uint32_t *p, uint32_t v) { __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }
static inline uint32_t at_take(uint32_t *p, uint32_t v) { return __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }
static inline void at_or(uint32_t *p, uint32_t v) { __atomic_fetch_or(p, v, __ATOMIC_RELAXED); }
static inline void at_max(uint32_t *p, uint32_t v) {
    uint32_t o = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (o < v && !__atomic_compare_exchange_n(p, &o, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
}
#else
#define PAR_FOR
#define PAR_FOR_MOST
static inline void at_add(uint32_t *p, uint32_t v) { *p += v; }
static inline uint32_t at_take(uint32_t *p, uint32_t v) { uint32_t o = *p; *p += v; return o; }
static inline void at_or(uint32_t *p, uint32_t v) { *p |= v; }
static inline void at_max(uint32_t *p, uint32_t v) { if (*p < v) *p = v; }
#endif

/* ------------------------------------------------------------ constants */
/* js/shared.js:270-289, fluid.wgsl:183-242,2062-2063,2221,2392,2579 */
#define BLOCK 8
#define LIQUID_CELL_WORDS 11
#define LIQUID_PER_CELL 4
#define FLUID_STATS 9
#define MG_COARSEST 512
#define MG_LEVELS 8
#define NONE 0xffffffffu
#define G_COUNT 0u
#define G_TAKEN 1u
#define G_REGION 1u
#define G_CAME 1u
#define G_SORT 2u
#define G_START 3u
#define G_DEST 4u
#define G_TYPE 5u
#define G_DENSITY 6u
#define G_MASS 7u
#define G_AIR_TO 8u
#define G_RIM 9u
#define FIX 65536.0f
#define MARK_LIQUID 1u
#define MARK_REGION 2u
#define ACTIVE 1u
#define HAS_FLUID 2u
#define HAS_VISCOUS 4u
#define NEAR_LIQUID 8u
#define COUNT_ACTIVE 0u
#define COUNT_RETIRED 1u
#define COUNT_VISCOUS 2u
#define COUNT_PARTICLES 3u
#define COUNT_RESIDUAL 4u
#define COUNT_CYCLES 7u
#define COUNT_MOST_CYCLES 8u
#define SMOOTH_OMEGA 1.0f
#define OMEGA 1.8f
#define SPREAD 0.3f
#define MAX_DRIFT 0.5f
#define DRIFT_KEEP 0.9f
#define PROJECTED 2.0f
#define AFFINE_KEEP 0.95f
#define LIQUID_MAX_SPEED 8.0f
#define ADVECT_STEPS 17
#define THETA_MIN 0.1f
#define RIM 10
#define PIN 1.0f
#define REGION_OPEN 0x80000000u
#define CAME_PRESSING 1u
#define CAME_FALLING 2u
#define APART (1.0f / 256.0f)
#define APART_FIX 16777216.0f
#define BUBBLE_GAS 0.25f
#define COLUMN_MOST 512
#define POWDER_PUSH 0.3f
#define RASTER_KINDS 4u
#define ESCAPE_REACH 16
#define SLIP 1.0f
#define EXACT_FINE 4294967296.0f
/* common.wgsl:139-190,262-270 */
#define DEST_KEEP 0xffffffffu
#define DEST_ASIDE 0x80000000u
#define DEST_CELL 0x40000000u
#define DEST_SPAN 0x3fffffffu
#define LIQUID_COUNT_MAX 255u
#define FL_DRAWN 4u
/* js/elements.js:8, element ids by table order (:424-457) */
#define KIND_SOLID 1u
#define KIND_POWDER 2u
#define KIND_LIQUID 3u
#define KIND_GAS 4u
#define E_AIR 0u
#define E_WATER 3u
/* water (js/elements.js:25-30,449-457,1122-1128) */
#define WATER_WEIGHT ((float)(9.81 / 3600.0 / 0.01))
#define WATER_HEATCAP 4.18f
#define WATER_LOT 0.0f
#define WATER_HIT 100.0f
#define WATER_UPSPAN ((float)(2257.0 / 4.18))
#define WATER_DOWNSPAN ((float)(334.0 / 4.18))
#define WATER_DOWNSCALE ((float)(4.18 / 1.9))
#define WATER_EXPANSION 1.42e-4f
#define WATER_EXPANSION2 4.45e-6f
#define AMBIENT_T 20.0f
#define KELVIN 273.15f
#define VAPOUR_MAX 8.0f
#define BOIL_K (100.0f + KELVIN)
/* the air (js/shared.js:58,137-175,238,289; common.wgsl:287-288;
 * js/elements.js:424,449-457,912-914: the packed table's f32 values) */
#define P_ATM 12.0f
#define AIR_BLOCKED (-10000.0f)
#define AIR_SOLID (-20000.0f)
#define WATER_RHO 1000.0f
#define AIR_RHO ((float)(1000.0 * 0.001204))
#define AIR_HEATCAP ((float)(1.2e-3 * (1000.0 / 8.0 / 0.598)))
#define VAPOUR_HEAT_CAP 2.0f
#define WATER_VAPOUR 64.0f
#define SOURCE_WORDS 16u
#define SOURCE_RIGHT 0u
#define SOURCE_DOWN 1u
#define SOURCE_AMOUNT 2u
#define SOURCE_VAPOUR 3u
#define SOURCE_MIST 4u
#define SOURCE_LOST 5u
#define SOURCE_HEAT 6u
/* the bodies (js/shared.js:294-333; common.wgsl:353,377-387; fluid.wgsl:1304-1318,2963) */
#define MAX_BODIES WIG_MAX_BODIES
#define BODY_TABLE_WORDS 116u
#define BODY_CONTACTS 16384u
#define BODY_CONTACT_WORDS 12u
#define BODY_LIQUID_WORDS 4u
#define BODY_PUSH_WORDS 36u
#define BW_POS 0u
#define BW_VEL 3u
#define BW_SPIN 5u
#define BW_MASS 6u
#define BW_INERTIA 7u
#define BW_CELLS 13u
#define BW_FLAGS 18u
#define BW_SUBCELL 25u
#define BW_CARGO 105u
#define BW_CARGO_INERTIA 106u
#define BW_MOBILITY 108u
#define BW_MOBILITY_AT 114u
#define BF_ALIVE 1u
#define BF_NEW 2u
#define BF_HELD 32u
#define BF_WET 64u
#define FL_LOOSE 8u
#define RIGID_FIX 16777216.0f
#define EXACT_COARSE 16777216.0f
#define LIQUID_WET 0x80000000u
#define CLAIM_PERIOD 0xfffffu
#define SEEP_REACH 128
#define PUSH_P0 0u
#define PUSH_X 2u
#define PUSH_F 3u
#define PUSH_SURFACE 4u
#define PUSH_DRAG 5u

/* ----------------------------------------------------------- the world */
typedef struct {
    double rz, alpha, beta, pq, rr, bb;
    uint32_t cycles, converged, pad[2];
} SolveHead;

typedef struct OWorld {
    int32_t W, H;
    uint32_t C, F, B, bw, bh, cap, L;
    uint32_t lw[MG_LEVELS + 1], lh[MG_LEVELS + 1];
    uint64_t lstart[MG_LEVELS + 2];
    uint64_t nfield;             /* vec4s */
    uint64_t bytes[WIG_PART_COUNT];
    uint32_t *cells;
    uint32_t *sorted, *parts;
    uint32_t *counters;
    float *field;
    uint32_t *blocks;
    float *cg;
    uint8_t *solve;              /* SolveHead, then 4 doubles a block */
    float *air;                  /* air[0]: a vec4 an air cell */
    uint32_t *source;            /* airSource: SOURCE_WORDS an air cell */
    uint32_t *occ;               /* occupants: occupantWord a cell */
    float *grav;                 /* the gravity pass's field: a vec4 a light cell */
    uint32_t *motion;            /* the bodies' table and liquid slots (WIG_PART_MOTION) */
    uint32_t *swept;             /* the bodies buffer's swept words, two a cell (WIG_PART_SWEPT) */
    uint32_t *pushes;            /* the step's push sums (WIG_PART_PUSHES) */
    uint8_t kind[256];
    WigElementProps props[256];  /* the elements' liquid columns */
    int dot_mode;
    /* the step being run */
    const WigParams *P;
    const WigStepCtl *ctl;
    uint32_t cycle;
    int converged_flag;
    /* captured at the solve's start for the FP64 oracle */
    float *cap_b;
    /* the solve's b (= -divergence at fluid cells), for residual replacement */
    float *bvec;
    int capture;
    uint32_t last_cycles;        /* iterations the last chunked solve ran */
} OWorld;

typedef struct { float x, y; } v2;
typedef struct { int32_t x, y; } i2;

static inline v2 V2(float x, float y) { v2 r = {x, y}; return r; }
static inline i2 I2(int32_t x, int32_t y) { i2 r = {x, y}; return r; }

/* ------------------------------------------------------ WGSL semantics */
static inline float wround(float x) { return rintf(x); }   /* ties to even */
static inline float wclamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline float wmix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
static inline float wlength(v2 v) { return sqrtf(v.x * v.x + v.y * v.y); }
static inline int32_t wi32(float x) {
    if (x != x) return 0;
    if (x <= -2147483648.0f) return INT32_MIN;
    if (x >= 2147483648.0f) return INT32_MAX;
    return (int32_t)x;
}
static inline uint32_t wu32(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)x;
}
static inline float fbits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t ubits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* IEEE binary16 <-> binary32, round to nearest even (pack2x16float). */
static uint16_t f16_of(float f) {
    uint32_t x = ubits(f);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);          /* rounds past 65504 */
    if (ax < 0x38800000u) {                                            /* subnormal half */
        if (ax < 0x33000000u) return (uint16_t)sign;                   /* < 2^-25 */
        uint32_t e = ax >> 23, m = (ax & 0x7fffffu) | 0x800000u;
        uint32_t shift = 126u - e;                                     /* to 2^-24 units: m * 2^(e-127-23+24) */
        uint32_t q = m >> (shift - 0u);
        uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1u);
        /* value = m * 2^(e-150); half units 2^-24 -> q = m >> (150-e-24) = m >> (126 - e) */
        if (rem > half || (rem == half && (q & 1u))) q++;
        return (uint16_t)(sign | q);
    }
    uint32_t e = (ax >> 23) - 112u, m = ax & 0x7fffffu;
    uint32_t h = (e << 10) | (m >> 13), rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}
static float f32_of(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 31u, m = h & 1023u;
    if (e == 31u) return fbits(sign | 0x7f800000u | (m << 13));
    if (e == 0u) {
        float v = (float)m * (1.0f / 16777216.0f);
        return sign ? -v : v;
    }
    return fbits(sign | ((e + 112u) << 23) | (m << 13));
}
static inline uint32_t pack2x16(v2 v) { return (uint32_t)f16_of(v.x) | ((uint32_t)f16_of(v.y) << 16); }
static inline v2 unpack2x16(uint32_t u) { return V2(f32_of((uint16_t)(u & 0xffffu)), f32_of((uint16_t)(u >> 16))); }

/* cell word accessors (js/shared.js:79,86-111,499-531) */
static inline uint32_t wType(uint32_t x) { return x & 0xffu; }
static inline uint32_t wAux(uint32_t x) { return (x >> 8) & 0xffu; }
static inline uint32_t wLife(uint32_t x) { return (x >> 16) & 0xffu; }
static inline uint32_t wShade(uint32_t x) { return (x >> 24) & 0xfu; }
static inline uint32_t wFlags(uint32_t x) { return x >> 28; }
static inline uint32_t packWord(uint32_t t, uint32_t aux, uint32_t life, uint32_t shade, uint32_t flags) {
    return (t & 0xffu) | ((aux & 0xffu) << 8) | ((life & 0xffu) << 16) | ((shade & 0xfu) << 24) | ((flags & 0xfu) << 28);
}
static inline uint32_t wWith(uint32_t x, uint32_t f) { return x | (f << 28); }
static inline uint32_t wWithout(uint32_t x, uint32_t f) { return x & ~(f << 28); }
static inline int wHas(uint32_t x, uint32_t f) { return (wFlags(x) & f) != 0u; }

/* common.wgsl:1413-1432 — the PCG hash (Jarzynski & Olano 2020) */
static inline uint32_t pcg(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28) + 4u)) ^ s) * 277803737u;
    return (w >> 22) ^ w;
}
typedef struct { uint32_t s; } Rng;
static inline void seedRng(Rng *r, uint32_t a, uint32_t b, uint32_t c) { r->s = pcg(a ^ pcg(b ^ pcg(c))); }
static inline uint32_t rndU(Rng *r) { r->s = pcg(r->s); return r->s; }
static inline float rnd(Rng *r) { return (float)(rndU(r) >> 8) * (1.0f / 16777216.0f); }

/* common.wgsl:1379-1411 — exact sums */
typedef struct { uint32_t lo, hi; } u2;
static u2 exactParts(float x, float scale) {
    float s = wround(wclamp(x * scale, -9.2e18f, 9.2e18f));
    u2 r;
    if (fabsf(s) < 2147483648.0f) {
        int32_t v = wi32(s);
        r.lo = (uint32_t)v;
        r.hi = v < 0 ? 0xffffffffu : 0u;
        return r;
    }
    float hi = floorf(s / 4294967296.0f);
    r.lo = wu32(s - hi * 4294967296.0f);
    r.hi = (uint32_t)wi32(hi);
    return r;
}
static float exactValue(u2 w, float scale) {
    int32_t l = (int32_t)w.lo, h = (int32_t)w.hi;
    if ((h == 0 && l >= 0) || (h == -1 && l < 0)) return (float)l / scale;
    return ((float)h * 4294967296.0f + (float)w.lo) / scale;
}
static u2 addExact(u2 w, float x) {                 /* fluid.wgsl:4127-4131 */
    u2 p = exactParts(x, EXACT_FINE);
    u2 r;
    r.lo = w.lo + p.lo;
    r.hi = w.hi + p.hi + (r.lo < p.lo ? 1u : 0u);
    return r;
}

/* common.wgsl:52-90 — the eight neighbour steps */
static uint32_t stepIndex(i2 d) {
    if (d.y == 0) return d.x > 0 ? 0u : 4u;
    if (d.x == 0) return d.y > 0 ? 2u : 6u;
    if (d.y > 0) return d.x > 0 ? 1u : 3u;
    return d.x > 0 ? 7u : 5u;
}
static int32_t stepX(uint32_t i) {
    uint32_t j = (i + 1u) & 7u;
    return j <= 2u ? 1 : ((j >= 4u && j <= 6u) ? -1 : 0);
}
static i2 stepAt(uint32_t i) { return I2(stepX(i), stepX((i + 6u) & 7u)); }
static i2 turnStep(i2 d, int32_t k) { return stepAt((stepIndex(d) + (uint32_t)(k + 8)) & 7u); }

/* ------------------------------------------------------- element table */
static inline uint32_t kindOf(const OWorld *w, uint32_t t) { return w->kind[t & 0xffu]; }
static inline int holdsSolute(uint32_t t) { return t == E_WATER; }

/* effTempIn at 1 atm with water's phase points (common.wgsl:1501-1530) */
static float effTemp(float stored) {
    return wclamp(stored, WATER_LOT, WATER_HIT) + fmaxf(stored - WATER_HIT - WATER_UPSPAN, 0.0f)
        - fmaxf(WATER_LOT - WATER_DOWNSPAN - stored, 0.0f) * WATER_DOWNSCALE;
}
/* expandedAt (common.wgsl:1560-1564) */
static float expandedAt(float temp) {
    float d = temp - fmaxf(AMBIENT_T, WATER_LOT);
    return wclamp(1.0f - WATER_EXPANSION * d - WATER_EXPANSION2 * d * d, 0.8f, 1.05f);
}
/* liquidDensity (common.wgsl:981-990): water's density 100 / 100 */
static float liquidDensity(uint32_t x, float temp) {
    (void)x;
    return 100.0f / 100.0f * expandedAt(effTemp(temp));
}
/* liquidVolume (fluid.wgsl:1048-1052) */
static float liquidVolume(uint32_t x, float temp) {
    (void)x;
    return 1.0f / expandedAt(effTemp(temp));
}

/* ------------------------------------------------------------- layout */
static uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1u) / b; }

static void layout(OWorld *w, uint32_t width, uint32_t height, uint32_t coarse) {
    w->W = (int32_t)width;
    w->H = (int32_t)height;
    w->C = width * height;
    w->F = (width + 1u) * height + width * (height + 1u);
    w->bw = ceil_div(width, BLOCK);
    w->bh = ceil_div(height, BLOCK);
    w->B = w->bw * w->bh;
    w->cap = (uint32_t)ceil((double)w->C * LIQUID_PER_CELL * 1.25);
    /* js/sim.js:763-770 */
    uint32_t L = 0;
    w->lw[0] = width;
    w->lh[0] = height;
    for (uint32_t l = 1; l <= MG_LEVELS; l++) {
        uint32_t sw = (width + (1u << l) - 1u) >> l, sh = (height + (1u << l) - 1u) >> l;
        w->lw[l] = sw;
        w->lh[l] = sh;
        L = l;
        if (coarse ? l >= coarse : sw * sh <= MG_COARSEST) break;
    }
    w->L = L;
    /* levelStart (fluid.wgsl:2237-2244) */
    uint64_t at = 2ull * w->C + w->F + 2ull * ((w->F + 3u) / 4u);
    for (uint32_t l = 1; l <= L + 1; l++) {
        w->lstart[l] = at;
        if (l <= L) at += 2ull * w->lw[l] * w->lh[l];
    }
    w->nfield = at;
    w->bytes[WIG_PART_CELLS] = 16ull * w->C + 16ull * ((w->C + 1u) / 2u);
    w->bytes[WIG_PART_PARTICLES0] = 16ull + 48ull * w->cap;
    w->bytes[WIG_PART_PARTICLES1] = w->bytes[WIG_PART_PARTICLES0];
    w->bytes[WIG_PART_COUNTERS] = 4ull * (LIQUID_CELL_WORDS * (uint64_t)w->C + 2ull * w->F);
    w->bytes[WIG_PART_FIELD] = 16ull * w->nfield;
    w->bytes[WIG_PART_BLOCKS] = 4ull * (5ull * w->B + FLUID_STATS);
    w->bytes[WIG_PART_FLUIDCG] = 16ull * (w->C + w->B + 1u);
    w->bytes[WIG_PART_SOLVE] = sizeof(SolveHead) + 32ull * w->B + 64u;   /* (and the library's 16 solver words) */
    w->bytes[WIG_PART_AIR] = 16ull * w->C;
    w->bytes[WIG_PART_AIR_SOURCE] = 4ull * SOURCE_WORDS * w->C;
    w->bytes[WIG_PART_OCCUPANTS] = 4ull * w->C;
    w->bytes[WIG_PART_GRAVITY] = 16ull * (uint64_t)((w->W + 1) / 2) * (uint64_t)((w->H + 1) / 2);
    w->bytes[WIG_PART_MOTION] = 4ull * WIG_MOTION_WORDS;
    w->bytes[WIG_PART_SWEPT] = 8ull * w->C;
    w->bytes[WIG_PART_PUSHES] = 4ull * WIG_MAX_BODIES * WIG_BODY_PUSH_WORDS;
}

/* ------------------------------------------------------------ address */
#define CNT(i, g) (w->counters[LIQUID_CELL_WORDS * (uint64_t)(i) + (g)])
#define SUM(f, k) (w->counters[LIQUID_CELL_WORDS * (uint64_t)w->C + 2ull * (f) + (k)])
#define FLD(v, c) (w->field[4ull * (uint64_t)(v) + (c)])
#define CG(i, c) (w->cg[4ull * (uint64_t)(i) + (c)])
#define MOTION(pb, k, c) ((pb)[4u + 4ull * (uint64_t)(k) + (c)])
#define STATE(pb, k, c) ((pb)[4u + 4ull * ((uint64_t)w->cap + (k)) + (c)])
#define AFFINE(pb, k, c) ((pb)[4u + 4ull * (2ull * w->cap + (k)) + (c)])

static inline SolveHead *head(OWorld *w) { return (SolveHead *)w->solve; }
static inline double *bpart(OWorld *w) { return (double *)(w->solve + sizeof(SolveHead)); }

static inline int inWorld(const OWorld *w, i2 c) { return c.x >= 0 && c.y >= 0 && c.x < w->W && c.y < w->H; }
static inline uint32_t cellIndex(const OWorld *w, i2 c) { return (uint32_t)c.y * (uint32_t)w->W + (uint32_t)c.x; }
static inline i2 cellOf(const OWorld *w, v2 p) {
    int32_t x = wi32(floorf(p.x)), y = wi32(floorf(p.y));
    x = x < 0 ? 0 : (x > w->W - 1 ? w->W - 1 : x);
    y = y < 0 ? 0 : (y > w->H - 1 ? w->H - 1 : y);
    return I2(x, y);
}
static inline i2 beside(i2 c, int d) {
    return d < 2 ? I2(c.x + (d & 1) * 2 - 1, c.y) : I2(c.x, c.y + (d & 1) * 2 - 1);
}
static inline uint32_t uFace(const OWorld *w, int32_t x, int32_t y) { return (uint32_t)y * (uint32_t)(w->W + 1) + (uint32_t)x; }
static inline uint32_t vFace(const OWorld *w, int32_t x, int32_t y) {
    return (uint32_t)(w->W + 1) * (uint32_t)w->H + (uint32_t)y * (uint32_t)w->W + (uint32_t)x;
}
static inline int faceInWorld(const OWorld *w, i2 at, int isU) {
    int32_t ex = w->W + (isU ? 1 : 0), ey = w->H + (isU ? 0 : 1);
    return at.x >= 0 && at.y >= 0 && at.x < ex && at.y < ey;
}
static inline uint32_t faceIndex(const OWorld *w, i2 at, int isU) { return isU ? uFace(w, at.x, at.y) : vFace(w, at.x, at.y); }
static inline uint64_t faceSlot(const OWorld *w, uint32_t f) { return 2ull * w->C + f; }
static inline uint64_t viscLastSlot(const OWorld *w, uint32_t f) { return 2ull * w->C + w->F + (w->F + 3u) / 4u + f / 4u; }

/* blocks (fluid.wgsl:444-513) */
static inline int inBlocks(const OWorld *w, i2 b) { return b.x >= 0 && b.y >= 0 && b.x < (int32_t)w->bw && b.y * (int32_t)w->bw < (int32_t)w->B; }
static inline uint32_t blockIndex(const OWorld *w, i2 b) { return (uint32_t)b.y * w->bw + (uint32_t)b.x; }
#define STATE_SLOT(n) (w->B + (n))
#define ACTIVE_SLOT(n) (2u * w->B + (n))
#define RETIRED_SLOT(n) (3u * w->B + (n))
#define COUNT_SLOT(k) (4u * w->B + (k))
#define TOTAL_SLOT(n) (4u * w->B + FLUID_STATS + (n))
static inline uint32_t activeCount(const OWorld *w) { return w->blocks[COUNT_SLOT(COUNT_ACTIVE)]; }
static inline int isActiveBlock(const OWorld *w, i2 b) { return inBlocks(w, b) && (w->blocks[STATE_SLOT(blockIndex(w, b))] & ACTIVE) != 0u; }
static inline i2 activeBlock(const OWorld *w, uint32_t n) {
    uint32_t b = w->blocks[ACTIVE_SLOT(n)];
    return I2((int32_t)(b % w->bw), (int32_t)(b / w->bw));
}
static inline i2 idiv8(i2 c) { return I2(c.x / BLOCK, c.y / BLOCK); }   /* WGSL integer division truncates */
static uint32_t statesOf(const OWorld *w, i2 b0, i2 b1) {
    uint32_t bits = 0;
    for (int32_t y = b0.y; y <= b1.y; y++)
        for (int32_t x = b0.x; x <= b1.x; x++)
            if (inBlocks(w, I2(x, y))) bits |= w->blocks[STATE_SLOT(blockIndex(w, I2(x, y)))];
    return bits;
}

/* facesOf (fluid.wgsl:296-303): (x, y, isU), isU < 0 none */
typedef struct { int32_t x, y, isU; } Face3;
static void facesOf(const OWorld *w, i2 c, Face3 fs[4]) {
    int right = c.x == w->W - 1 || (c.x % BLOCK == BLOCK - 1 && !isActiveBlock(w, I2(c.x / BLOCK + 1, c.y / BLOCK)));
    int bottom = c.y == w->H - 1 || (c.y % BLOCK == BLOCK - 1 && !isActiveBlock(w, I2(c.x / BLOCK, c.y / BLOCK + 1)));
    fs[0] = (Face3){c.x, c.y, 1};
    fs[1] = (Face3){c.x, c.y, 0};
    fs[2] = (Face3){c.x + 1, c.y, right ? 1 : -1};
    fs[3] = (Face3){c.x, c.y + 1, bottom ? 0 : -1};
}

/* the gas plane (common.wgsl:412-465) */
static inline uint64_t gasWord(const OWorld *w, uint32_t i) { return 4ull * (w->C + (i >> 1)) + ((i & 1u) ? 2u : 0u); }
static inline float gasAmountAt(const OWorld *w, uint32_t i) {
    uint64_t at = gasWord(w, i);
    v2 m = unpack2x16(w->cells[at]);
    v2 r = unpack2x16(w->cells[at + 1]);
    return r.y + m.x * (AMBIENT_T + KELVIN) / (BOIL_K * VAPOUR_MAX);
}
static inline void setGasPair(OWorld *w, uint32_t i, uint32_t a, uint32_t b) {
    uint64_t at = gasWord(w, i);
    w->cells[at] = a;
    w->cells[at + 1] = b;
}

/* --------------------------------------------------------- cell tests */
static inline float densityAt(const OWorld *w, i2 c) {       /* :322-327 */
    if (!inWorld(w, c)) return 0.0f;
    return (float)CNT(cellIndex(w, c), G_DENSITY) / (FIX * (float)LIQUID_PER_CELL);
}
static inline float gridZ(const OWorld *w, i2 c) { return inWorld(w, c) ? FLD(2u * cellIndex(w, c), 2) : -1.0f; }
static inline float gridX(const OWorld *w, i2 c) { return inWorld(w, c) ? FLD(2u * cellIndex(w, c), 0) : 0.0f; }
static inline float driftAt(const OWorld *w, i2 c) { return inWorld(w, c) ? FLD(2u * cellIndex(w, c) + 1u, 0) : 0.0f; }
static inline float faceBeta(float a, float b) {             /* :369-374 */
    if (a < 0.0f || b < 0.0f || (a <= 0.0f && b <= 0.0f)) return 0.0f;
    return (a > 0.0f && b > 0.0f) ? 2.0f / (a + b) : 1.0f / fmaxf(a, b);
}
static inline float surfaceShare(float da, float db) {      /* :384-386 */
    return wclamp((da - 0.5f) / fmaxf(da - db, 1e-3f), THETA_MIN, 1.0f);
}
static inline float fineBeta(float a, float b, float da, float db) {   /* :396-405 */
    float beta = faceBeta(a, b);
    if (a > 0.0f && b == 0.0f) return beta / (da > 0.0f ? surfaceShare(da, db) : 0.5f);
    if (b > 0.0f && a == 0.0f) return beta / (db > 0.0f ? surfaceShare(db, da) : 0.5f);
    return beta;
}
static inline int isLiquidAt(const OWorld *w, i2 c) {
    return inWorld(w, c) && kindOf(w, wType(w->cells[4ull * cellIndex(w, c)])) == KIND_LIQUID;
}
static inline uint32_t stepBit(const OWorld *w, uint32_t back) { return 1u << ((w->P->frame + back) & 1u); }
static inline uint32_t fallingBit(const OWorld *w, uint32_t back) { return 4u << ((w->P->frame + back) & 1u); }
static inline int fallingAt(const OWorld *w, i2 q) {         /* :1120-1122 */
    return inWorld(w, q) && (wu32(FLD(2u * cellIndex(w, q) + 1u, 3)) & fallingBit(w, 1u)) != 0u;
}
/* people's bodies (common.wgsl:238-254; fluid.wgsl:175-178): cell i's
 * occupantWord, 0 for nobody or with the occupants off */
#define OCCUPANT_SCALE 128.0f
static inline uint32_t occupantAt(const OWorld *w, uint32_t i) { return w->ctl->occupants ? w->occ[i] : 0u; }
static inline int isOccupant(uint32_t o) { return (o & 0xffu) != 0u; }
static inline v2 occupantVel(uint32_t o) {
    /* the 12-bit fields sign-extended by shifting them to the top first */
    int32_t x = (int32_t)(o << 12u) >> 20, y = (int32_t)o >> 20;
    if (MUTANT(28)) y = (int32_t)(o << 12u) >> 20;
    return V2((float)x / OCCUPANT_SCALE, (float)y / OCCUPANT_SCALE);
}
/* The pull and down (pull.wgsl:13-44) from the gravity pass's field, a vec4
 * a light cell of GRAVITY_BLOCK x GRAVITY_BLOCK cells: the masses' pull
 * (xy), their potential (z), down (w).  Off, Earth's pull alone and down
 * (0, 1): the field's texel for that, no masses' pull and down step 2. */
#define GRAVITY_BLOCK 2
static void gravityTexel(const OWorld *w, i2 c, float t[4]) {
    if (!w->ctl->gravity) { t[0] = t[1] = t[2] = 0.0f; t[3] = 2.0f; return; }
    int32_t gw = (w->W + 1) / 2, gh = (w->H + 1) / 2;
    int32_t x = c.x < 0 ? 0 : (c.x > gw - 1 ? gw - 1 : c.x), y = c.y < 0 ? 0 : (c.y > gh - 1 ? gh - 1 : c.y);
    memcpy(t, &w->grav[4ull * ((uint64_t)y * (uint64_t)gw + (uint64_t)x)], 16);
}
static v2 fieldPull(const OWorld *w, v2 pos) {                /* fieldAt's xy (pull.wgsl:19-30) */
    float qx = pos.x / (float)GRAVITY_BLOCK - 0.5f, qy = pos.y / (float)GRAVITY_BLOCK - 0.5f;
    float fx0 = floorf(qx), fy0 = floorf(qy);
    i2 i = I2(wi32(fx0), wi32(fy0));
    float fx = qx - fx0, fy = qy - fy0;
    float a[4], b[4], c[4], e[4];
    gravityTexel(w, i, a);
    gravityTexel(w, I2(i.x + 1, i.y), b);
    gravityTexel(w, I2(i.x, i.y + 1), c);
    gravityTexel(w, I2(i.x + 1, i.y + 1), e);
    if (MUTANT(30)) fx = 1.0f - fx;
    v2 top = V2(wmix(a[0], b[0], fx), wmix(a[1], b[1], fx)), bottom = V2(wmix(c[0], e[0], fx), wmix(c[1], e[1], fx));
    return V2(wmix(top.x, bottom.x, fy), wmix(top.y, bottom.y, fy));
}
static v2 pullAt(const OWorld *w, v2 pos) {                   /* pull.wgsl:32-34 */
    if (!w->ctl->gravity) return V2(0.0f, w->P->earthG);
    v2 f = fieldPull(w, pos);
    return V2(0.0f + f.x, w->P->earthG + f.y);
}
static i2 downAt(const OWorld *w, i2 p) {                     /* pull.wgsl:36-44, p in the world */
    if (!w->ctl->gravity) return I2(0, 1);
    float t[4];
    gravityTexel(w, I2(p.x / GRAVITY_BLOCK, p.y / GRAVITY_BLOCK), t);
    return stepAt(wu32(t[3]) & 7u);
}
static int32_t downTurns(const OWorld *w, i2 c) {             /* fluid.wgsl:3822-3825 */
    i2 at = I2(c.x < 0 ? 0 : (c.x > w->W - 1 ? w->W - 1 : c.x), c.y < 0 ? 0 : (c.y > w->H - 1 ? w->H - 1 : c.y));
    if (MUTANT(31)) return 0;
    return (int32_t)stepIndex(downAt(w, at)) - 2;
}
static int isObstacleCell(const OWorld *w, i2 c, uint32_t vx) {   /* :1123-1147 */
    if (inWorld(w, c) && isOccupant(occupantAt(w, cellIndex(w, c)))) return 1;
    uint32_t t = wType(vx), k = kindOf(w, t);
    if (k == KIND_GAS) {
        int wet = 0;
        for (int d = 0; d < 4; d++) wet += isLiquidAt(w, beside(c, d)) ? 1 : 0;
        if (wet < 2) return 0;
        if (t == E_AIR) {
            i2 dn = downAt(w, c);
            i2 up = I2(c.x - dn.x, c.y - dn.y), up2 = I2(c.x - 2 * dn.x, c.y - 2 * dn.y);   /* c - downAt(c) */
            if (!isLiquidAt(w, up) || !isLiquidAt(w, up2) || fallingAt(w, up)) return 0;
            return gasAmountAt(w, cellIndex(w, c)) >= BUBBLE_GAS;
        }
        return 1;
    }
    return k == KIND_SOLID || k == KIND_POWDER;
}
static int32_t countAt(const OWorld *w, i2 c) {              /* :1150-1159 */
    if (!inWorld(w, c)) return -1;
    uint32_t i = cellIndex(w, c);
    if (isObstacleCell(w, c, w->cells[4ull * i])) return -1;
    return (int32_t)CNT(i, G_COUNT);
}
static void gridAt4(const OWorld *w, i2 c, float g[4]) {
    if (!inWorld(w, c)) { g[0] = 0.0f; g[1] = 0.0f; g[2] = -1.0f; g[3] = 0.0f; return; }
    for (int j = 0; j < 4; j++) g[j] = FLD(2u * cellIndex(w, c), j);
}
/* ------------------------------------------------------------ bodies
 * Free and held bodies in the liquid (fluid.wgsl:1181-1475): the bodies'
 * table in `motion` (bodies.wgsl's, js/shared.js:294-333's BODY_WORD
 * words), the swept words of the bodies buffer (common.wgsl:355-370), and
 * the step's push sums (`pushes`).  With ctl->bodies off none is read and
 * no cell is a body's. */
typedef struct { float x, y, z; } v3;
static inline v3 V3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline int bodiesOn(const OWorld *w) { return w->ctl->bodies != 0u; }
static inline uint32_t bodyWord(const OWorld *w, uint32_t b, uint32_t k) {   /* :1211-1213 */
    return w->motion[MAX_BODIES + b * BODY_TABLE_WORDS + k];
}
static inline float bodyF(const OWorld *w, uint32_t b, uint32_t k) { return fbits(bodyWord(w, b, k)); }
static inline v2 bodyV(const OWorld *w, uint32_t b, uint32_t k) { return V2(bodyF(w, b, k), bodyF(w, b, k + 1u)); }
static inline uint32_t liquidSlot(uint32_t b, uint32_t wd) {                    /* :1226-1228 */
    return MAX_BODIES * (1u + BODY_TABLE_WORDS) + BODY_CONTACTS * BODY_CONTACT_WORDS + b * BODY_LIQUID_WORDS + wd;
}
/* alive and weighed: not found this step, its mass set (:1244-1246) */
static inline int bodyWeighed(const OWorld *w, uint32_t b) {
    uint32_t flags = bodyWord(w, b, BW_FLAGS);
    return !((flags & BF_ALIVE) == 0u || (flags & BF_NEW) != 0u || bodyF(w, b, BW_MASS) <= 0.0f);
}
static uint32_t rigidBodyOf(const OWorld *w, const uint32_t *v) {            /* :1233-1248 */
    if (!bodiesOn(w)) return NONE;
    uint32_t t = wType(v[0]);
    if (!wHas(v[0], FL_LOOSE) || kindOf(w, t) != KIND_SOLID || w->props[t].breaks == 0.0f) return NONE;
    float id = unpack2x16(v[3]).x;
    if (id < 0.0f || id >= (float)MAX_BODIES) return NONE;
    uint32_t b = wu32(id);
    return bodyWeighed(w, b) ? b : NONE;
}
static inline uint32_t claimStep(uint32_t frame) { return frame % CLAIM_PERIOD + 1u; }   /* common.wgsl:384-386 */
/* The body whose cup (seam 0) or seam (1) cell i was stamped as `back`
 * steps ago (common.wgsl:364-366's two words), alive and weighed when
 * `weighed` (cupOf :1252-1266, seamOf :1272-1286; cupBefore :3520-3527
 * asks only for a slot). */
static uint32_t sweptBody(const OWorld *w, uint32_t i, uint32_t back, uint32_t seam, int weighed) {
    if (!bodiesOn(w)) return NONE;
    uint32_t frame = w->P->frame - back;
    uint32_t key = w->swept[2ull * i + (frame & 1u)];
    if ((key >> 12) != claimStep(frame) || (key & 1u) != seam) return NONE;
    uint32_t b = 2047u - ((key >> 1) & 0x7ffu);
    if (b >= MAX_BODIES) return NONE;
    if (weighed && !bodyWeighed(w, b)) return NONE;
    return b;
}
static inline uint32_t cupOf(const OWorld *w, uint32_t i) { return sweptBody(w, i, 1u, 0u, 1); }
static inline uint32_t seamOf(const OWorld *w, uint32_t i) { return sweptBody(w, i, 1u, 1u, 1); }
static inline uint32_t cupBefore(const OWorld *w, uint32_t i) { return sweptBody(w, i, 2u, 0u, 0); }
static inline float rigidMass(const OWorld *w, uint32_t b) { return bodyF(w, b, BW_MASS) + bodyF(w, b, BW_CARGO); }   /* :1290-1296 */
static inline float rigidInertia(const OWorld *w, uint32_t b) { return bodyF(w, b, BW_INERTIA) + bodyF(w, b, BW_CARGO_INERTIA); }
static inline float bodyDensity(const OWorld *w, uint32_t b) {                  /* :1300-1302 */
    return rigidMass(w, b) / fmaxf((float)bodyWord(w, b, BW_CELLS), 1.0f);
}
static inline uint32_t pushSlot(uint32_t b, uint32_t which, uint32_t k) { return b * BODY_PUSH_WORDS + 6u * which + 2u * k; }
static void addPush(OWorld *w, uint32_t b, uint32_t which, v3 v) {             /* :1324-1331 */
    float c[3] = {v.x, v.y, v.z};
    for (uint32_t k = 0; k < 3u; k++) {
        u2 e = exactParts(c[k], EXACT_COARSE);
        uint32_t at = pushSlot(b, which, k);
        uint32_t old = at_take(&w->pushes[at], e.lo);
        at_add(&w->pushes[at + 1u], e.hi + (old + e.lo < old ? 1u : 0u));
    }
}
static v3 pushOf(const OWorld *w, uint32_t b, uint32_t which) {               /* :1333-1340 */
    float c[3];
    for (uint32_t k = 0; k < 3u; k++) {
        uint32_t at = pushSlot(b, which, k);
        u2 e = {w->pushes[at], w->pushes[at + 1u]};
        c[k] = exactValue(e, EXACT_COARSE);
    }
    return V3(c[0], c[1], c[2]);
}
static v3 giveWay(const OWorld *w, uint32_t b, v3 g) {                        /* :1349-1356 */
    if (bodyWord(w, b, BW_MOBILITY_AT) != w->P->frame - 1u)
        return V3(g.x / rigidMass(w, b), g.y / rigidMass(w, b), MUTANT(35) ? g.z / rigidMass(w, b) : g.z / rigidInertia(w, b));
    float m[6];
    for (uint32_t k = 0; k < 6u; k++) m[k] = bodyF(w, b, BW_MOBILITY + k);
    return V3(m[0] * g.x + m[1] * g.y + m[2] * g.z, m[1] * g.x + m[3] * g.y + m[4] * g.z, m[2] * g.x + m[4] * g.y + m[5] * g.z);
}
static inline v3 pushChange(const OWorld *w, uint32_t b, uint32_t which) { return giveWay(w, b, pushOf(w, b, which)); }
static inline v3 pushChangeFree(const OWorld *w, uint32_t b, uint32_t which) {   /* :1367-1370 */
    v3 g = pushOf(w, b, which);
    return V3(g.x / rigidMass(w, b), g.y / rigidMass(w, b), g.z / rigidInertia(w, b));
}
/* the face from cell c to body b's cell q: (n, r x n) (:1374-1378) */
static v3 bodyFaceJ(const OWorld *w, i2 c, i2 q, uint32_t b) {
    v2 n = V2((float)(q.x - c.x), (float)(q.y - c.y));
    v2 pos = bodyV(w, b, BW_POS);
    v2 r = V2((float)c.x + 0.5f + 0.5f * n.x - pos.x, (float)c.y + 0.5f + 0.5f * n.y - pos.y);
    return V3(n.x, n.y, MUTANT(36) ? r.y * n.x - r.x * n.y : r.x * n.y - r.y * n.x);
}
/* the body whose cell c is in the solve, as classify found it (:1422-1429) */
static uint32_t rigidAt(const OWorld *w, i2 c) {
    if (!bodiesOn(w) || !inWorld(w, c)) return NONE;
    float z = FLD(2u * cellIndex(w, c) + 1u, 2);
    uint32_t b = wu32(fmaxf(-z, 1.0f)) - 1u;
    return z < 0.0f && b < MAX_BODIES ? b : NONE;
}
static uint32_t heldAt(const OWorld *w, i2 c) {                               /* :1450-1457 */
    if (!bodiesOn(w) || !inWorld(w, c)) return NONE;
    float z = FLD(2u * cellIndex(w, c) + 1u, 2);
    uint32_t b = wu32(fmaxf(-z, 1.0f)) - 1u;
    return z < 0.0f && b >= MAX_BODIES ? b - MAX_BODIES : NONE;
}
static inline int isHeld(const OWorld *w, uint32_t b) { return (bodyWord(w, b, BW_FLAGS) & BF_HELD) != 0u; }
static inline v2 rigidVel(const OWorld *w, uint32_t b, v2 x) {                  /* :1460-1464 */
    v2 pos = bodyV(w, b, BW_POS), vel = bodyV(w, b, BW_VEL);
    v2 r = V2(x.x - pos.x, x.y - pos.y);
    float spin = bodyF(w, b, BW_SPIN);
    return V2(vel.x + -spin * r.y, vel.y + spin * r.x);
}
static inline int shutAt(const OWorld *w, i2 c) { return gridZ(w, c) < 0.0f || rigidAt(w, c) != NONE; }   /* :1468-1470 */
static inline v2 faceMiddle(i2 at, int isU) { return isU ? V2((float)at.x, (float)at.y + 0.5f) : V2((float)at.x + 0.5f, (float)at.y); }
static inline float bodyReach(const OWorld *w, uint32_t b, i2 n) {             /* :413-415 */
    v2 s = unpack2x16(bodyWord(w, b, BW_SUBCELL));
    return wclamp(0.5f + (s.x * (float)n.x + s.y * (float)n.y), 0.0f, 1.0f);
}
static void addPressurePush(OWorld *w, i2 c, float p, uint32_t which) {       /* :1381-1389 */
    for (int d = 0; d < 4; d++) {
        i2 q = beside(c, d);
        uint32_t b = rigidAt(w, q);
        if (b != NONE) {
            v3 j = bodyFaceJ(w, c, q, b);
            addPush(w, b, which, V3(j.x * p, j.y * p, j.z * p));
        }
    }
}
static float couplingAt(const OWorld *w, i2 c, uint32_t which) {             /* :1393-1403 */
    float sum = 0.0f;
    for (int d = 0; d < 4; d++) {
        i2 q = beside(c, d);
        uint32_t b = rigidAt(w, q);
        if (b != NONE) {
            v3 j = bodyFaceJ(w, c, q, b), ch = pushChange(w, b, which);
            sum += j.x * ch.x + j.y * ch.y + j.z * ch.z;
        }
    }
    return sum;
}
/* body b's velocity at x going into the solve (:1410-1419) */
static v2 bodyMoving(const OWorld *w, uint32_t b, v2 x, int drag) {
    v2 pos = bodyV(w, b, BW_POS);
    v2 pull = pullAt(w, pos);
    float mass = rigidMass(w, b);
    v3 s = pushChange(w, b, PUSH_SURFACE), g = giveWay(w, b, V3(mass * pull.x, mass * pull.y, 0.0f));
    v3 ch = V3(s.x + g.x, s.y + g.y, s.z + g.z);
    if (drag) {
        v3 dr = pushChange(w, b, PUSH_DRAG);
        ch = V3(ch.x + dr.x, ch.y + dr.y, ch.z + dr.z);
    }
    v2 r = V2(x.x - pos.x, x.y - pos.y);
    v2 rv = rigidVel(w, b, x), px = pullAt(w, x);
    return V2(rv.x + px.x - pull.x + ch.x + -ch.z * r.y, rv.y + px.y - pull.y + ch.y + ch.z * r.x);
}
/* the pull's potential at pos, Earth's and the masses' (:2954-2956) */
static float potentialAt(const OWorld *w, v2 pos) {
    float z = 0.0f;
    if (w->ctl->gravity) {
        float qx = pos.x / (float)GRAVITY_BLOCK - 0.5f, qy = pos.y / (float)GRAVITY_BLOCK - 0.5f;
        float fx0 = floorf(qx), fy0 = floorf(qy);
        i2 i = I2(wi32(fx0), wi32(fy0));
        float fx = qx - fx0, fy = qy - fy0;
        float a[4], b[4], c[4], e[4];
        gravityTexel(w, i, a);
        gravityTexel(w, I2(i.x + 1, i.y), b);
        gravityTexel(w, I2(i.x, i.y + 1), c);
        gravityTexel(w, I2(i.x + 1, i.y + 1), e);
        z = wmix(wmix(a[2], b[2], fx), wmix(c[2], e[2], fx), fy);
    }
    return w->P->earthG * ((float)w->P->size[1] - pos.y) + z;
}
/* the liquid's pressure seeping in at x on body h's face from c toward n
 * (:2976-3001) */
static float seepAt(const OWorld *w, i2 c, i2 n, uint32_t h, int held, v2 x) {
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
            i2 q = I2(c.x + ways[k2].x * k, c.y + ways[k2].y * k);
            if ((held ? heldAt(w, q) : rigidAt(w, q)) == h) continue;
            float g[4];
            gridAt4(w, q, g);
            if (g[2] > 0.0f) {
                nearest = k;
                p = g[0] + g[2] * (potentialAt(w, V2((float)q.x + 0.5f, (float)q.y + 0.5f)) - potentialAt(w, x));
            }
            break;
        }
    }
    return p;
}
/* what the liquid's pressure on held body h's cell c does to it (:3010-3035) */
static void heldPush(const OWorld *w, i2 c, uint32_t h, int32_t out[3]) {
    v2 pos = bodyV(w, h, BW_POS);
    v2 f = V2(0.0f, 0.0f);
    float torque = 0.0f;
    for (int d = 0; d < 4; d++) {
        i2 q = beside(c, d);
        i2 n = I2(q.x - c.x, q.y - c.y);
        if (heldAt(w, q) == h) continue;
        float g[4];
        gridAt4(w, q, g);
        v2 x = V2((float)c.x + 0.5f + 0.5f * (float)n.x, (float)c.y + 0.5f + 0.5f * (float)n.y);
        float p = g[0];
        if (g[2] > 0.0f) p += g[2] * (potentialAt(w, V2((float)q.x + 0.5f, (float)q.y + 0.5f)) - potentialAt(w, x));
        else if (g[2] < 0.0f) p = seepAt(w, c, n, h, 1, x);
        v2 push = V2(-p * (float)n.x, -p * (float)n.y);
        f = V2(f.x + push.x, f.y + push.y);
        v2 r = V2(x.x - pos.x, x.y - pos.y);
        torque += r.x * push.y - r.y * push.x;
    }
    float m = rigidMass(w, h), in = rigidInertia(w, h);
    out[0] = wi32(wround(f.x / m * RIGID_FIX));
    out[1] = wi32(wround(f.y / m * RIGID_FIX));
    out[2] = wi32(wround(torque / in * RIGID_FIX));
}
/* e^x in float from basic operations only (round to nearest, no fused
 * multiply-add), the same steps on the GPU, so both give the same bits:
 * x = k ln2 + r with k = rint(x / ln2) and ln2 split in two (Cody and
 * Waite), e^r by its Taylor series to r^7 on |r| <= ln2 / 2 (relative
 * error under 2^-23 there), times 2^k.  viscosityAt's exp. */
static float wexpf(float x) {
    if (!(x < 88.5f)) return x != x ? x : INFINITY;
    if (x < -103.0f) return 0.0f;
    float k = rintf(x * 1.44269504f);
    float r = (x - k * 0.693145752f) - k * 1.42860677e-6f;
    float p = 1.0f + r * (1.0f + r * (0.5f + r * (0.166666672f + r * (0.0416666679f + r * (0.00833333377f
              + r * (0.00138888892f + r * 0.000198412701f))))));
    return ldexpf(p, (int)k);
}
/* viscosityAt (common.wgsl:1486-1493): a liquid's viscosity at temp */
#define VISCOSITY_MOST 1e4f
static float viscosityAt(const WigElementProps *e, float temp) {
    if (e->viscosity_e <= 0.0f) return e->viscosity;
    return fminf(e->viscosity * wexpf((e->viscosity_t - temp) / e->viscosity_e), VISCOSITY_MOST);
}
static v2 liquidIn(const OWorld *w, uint32_t i) {           /* :1165-1179 */
    uint32_t vx = w->cells[4ull * i];
    float temp = fbits(w->cells[4ull * i + 1]);
    uint32_t k = kindOf(w, wType(vx));
    /* at the temperature it acts at: water's effTempIn (every liquid has
     * water's phase points here) */
    float viscosity = k == KIND_LIQUID ? viscosityAt(&w->props[wType(vx)], effTemp(temp)) : 0.0f;
    uint32_t n = CNT(i, G_COUNT);
    if (n > 0u) return V2((float)CNT(i, G_MASS) / (FIX * (float)n), viscosity);
    if (k != KIND_LIQUID) return V2(1.0f, 0.0f);
    return V2(liquidDensity(vx, temp), viscosity);
}
static int isObstaclePos(const OWorld *w, v2 pos) {          /* :424-431 */
    i2 c = I2(wi32(floorf(pos.x)), wi32(floorf(pos.y)));
    if (gridZ(w, c) < 0.0f || isOccupant(occupantAt(w, cellIndex(w, c)))) return 1;
    uint32_t k = kindOf(w, wType(w->cells[4ull * cellIndex(w, c)]));
    return k == KIND_SOLID || k == KIND_POWDER;
}
static inline int walls(const OWorld *w, i2 c) {            /* :1056-1062 */
    if (!inWorld(w, c) || isOccupant(occupantAt(w, cellIndex(w, c)))) return 1;
    uint32_t k = kindOf(w, wType(w->cells[4ull * cellIndex(w, c)]));
    return k == KIND_SOLID || k == KIND_POWDER;
}

/* ------------------------------------------------------------- the air
 * air[0] and airSource, as the liquid reads and adds to them (AIR_CELL 1:
 * an air cell is a world cell, airDims the world's size). */
static inline const float *airAt(const OWorld *w, i2 i) { return &w->air[4ull * ((uint64_t)i.y * (uint32_t)w->W + (uint32_t)i.x)]; }
static inline int inAir(const OWorld *w, i2 i) { return i.x >= 0 && i.y >= 0 && i.x < w->W && i.y < w->H; }
static v2 fluidPressureAt(const OWorld *w, i2 i) {          /* airsample.wgsl:83-89 */
    if (!inAir(w, i)) return V2(0.0f, 0.0f);
    const float *a = airAt(w, i);
    return V2(a[0], a[3] <= AIR_SOLID ? 0.0f : 1.0f);
}
static float pressureSlope(const OWorld *w, i2 c, i2 d) {   /* airsample.wgsl:97-111 */
    float own = airAt(w, c)[0];
    v2 lo = fluidPressureAt(w, I2(c.x - d.x, c.y - d.y)), hi = fluidPressureAt(w, I2(c.x + d.x, c.y + d.y));
    if (lo.y > 0.0f && hi.y > 0.0f) return 0.5f * (hi.x - lo.x);
    if (hi.y > 0.0f) return hi.x - own;
    if (lo.y > 0.0f) return own - lo.x;
    return 0.0f;
}
static v2 airGrad(const OWorld *w, v2 pos) {                /* sampleAir's grad, airsample.wgsl:171-180 */
    v2 q = V2(pos.x / 1.0f, pos.y / 1.0f);
    i2 c = I2(wi32(floorf(q.x)), wi32(floorf(q.y)));
    c.x = c.x < 0 ? 0 : (c.x > w->W - 1 ? w->W - 1 : c.x);
    c.y = c.y < 0 ? 0 : (c.y > w->H - 1 ? w->H - 1 : c.y);
    return V2(pressureSlope(w, c, I2(1, 0)) / 1.0f, pressureSlope(w, c, I2(0, 1)) / 1.0f);
}
static v2 pressurePush(const OWorld *w, v2 grad, float rho) {   /* airsample.wgsl:11-14 */
    float s = w->P->dt / w->P->dx, m = fmaxf(rho, 1e-3f);
    if (MUTANT(25)) return V2(grad.x * (101325.0f / P_ATM) * s * s / m, grad.y * (101325.0f / P_ATM) * s * s / m);
    return V2(-grad.x * (101325.0f / P_ATM) * s * s / m, -grad.y * (101325.0f / P_ATM) * s * s / m);
}
static inline int gasCellAt(const OWorld *w, i2 i) { return inAir(w, i) && airAt(w, i)[3] > AIR_BLOCKED; }   /* airsample.wgsl:130-132 */
static inline float faceU(const OWorld *w, i2 i) { return inAir(w, i) ? airAt(w, i)[1] : 0.0f; }           /* :115-127 */
static inline float faceV(const OWorld *w, i2 i) { return inAir(w, i) ? airAt(w, i)[2] : 0.0f; }
static void addSource(OWorld *w, uint32_t i, uint32_t k, float x) {   /* airsource.wgsl:27-35 */
    if (x == 0.0f) return;
    uint64_t at = SOURCE_WORDS * (uint64_t)i + 2u * k;
    u2 p = exactParts(x, EXACT_FINE);
    uint32_t old = at_take(&w->source[at], p.lo);
    at_add(&w->source[at + 1u], p.hi + (old + p.lo < old ? 1u : 0u));
}
static void pushFace(OWorld *w, i2 i, v2 dv) {               /* airsource.wgsl:37-41,59-63 */
    if (!inAir(w, i)) return;
    uint32_t j = (uint32_t)i.y * (uint32_t)w->W + (uint32_t)i.x;
    addSource(w, j, SOURCE_RIGHT, dv.x);
    addSource(w, j, SOURCE_DOWN, dv.y);
}
static void pushGasOut(OWorld *w, uint32_t i, float temp) {  /* fluid.wgsl:3855-3861; airsource.wgsl:65-72 */
    uint64_t at = gasWord(w, i);
    v2 m = unpack2x16(w->cells[at]), r = unpack2x16(w->cells[at + 1]);
    float vapour = m.x, mist = m.y, lost = r.x, amount = r.y;
    addSource(w, i, SOURCE_AMOUNT, amount);
    if (!MUTANT(27)) addSource(w, i, SOURCE_VAPOUR, vapour);
    addSource(w, i, SOURCE_MIST, mist);
    addSource(w, i, SOURCE_LOST, lost * amount);
    /* gasHeatCap (common.wgsl:473-475) times the temperature */
    addSource(w, i, SOURCE_HEAT, (AIR_HEATCAP * amount + (VAPOUR_HEAT_CAP * vapour + WATER_HEATCAP * mist) / WATER_VAPOUR) * temp);
    setGasPair(w, i, 0u, 0u);
}
static v2 pairDrag(float u, float wv, float ml, float ma, float c) {   /* fluid.wgsl:1869-1878 */
    if (ml <= 0.0f || ma <= 0.0f || c <= 0.0f) return V2(0.0f, 0.0f);
    float mu = ml / ma;
    float k = c * fabsf(u - wv);
    float change = (u - wv) * (1.0f - 1.0f / (1.0f + k * (1.0f + 1.0f / mu)));
    if (MUTANT(26)) return V2(-change / (1.0f + mu), change / (1.0f + mu));
    return V2(-change / (1.0f + mu), change * mu / (1.0f + mu));
}
static float airMassAt(const OWorld *w, i2 c) {             /* fluid.wgsl:1862-1865 */
    return AIR_RHO / 1000.0f * gasAmountAt(w, cellIndex(w, c));
}
static float surfaceDrag(OWorld *w, i2 a, i2 b, int isU, const float ga[4], const float gb[4], float u) {   /* :1914-1932 */
    i2 across = isU ? I2(0, 1) : I2(1, 0);
    float du = 0.0f;
    if (ga[2] > 0.0f && gb[2] > 0.0f) {
        float k = w->ctl->surface_drag;
        for (int s = -1; s <= 1; s += 2) {
            i2 oa = I2(a.x + across.x * s, a.y + across.y * s), ob = I2(b.x + across.x * s, b.y + across.y * s);
            if (gasCellAt(w, oa) && gasCellAt(w, ob)) {
                float wv = isU ? faceU(w, oa) : faceV(w, oa);
                v2 drag = pairDrag(u + du, wv, 0.5f * (ga[2] + gb[2]), 0.5f * (airMassAt(w, oa) + airMassAt(w, ob)), k);
                du += drag.x;
                pushFace(w, oa, isU ? V2(drag.y, 0.0f) : V2(0.0f, drag.y));
            }
        }
    }
    return du;
}

/* union-find (label.wgsl:25-55) over G_REGION words */
static inline uint32_t ufLink(OWorld *w, uint32_t at) { return CNT(at, G_REGION); }
static uint32_t ufFind(OWorld *w, uint32_t i) {
    uint32_t x = i;
    for (;;) {
        uint32_t up = ufLink(w, x);
        if (up >= x) break;
        x = up;
    }
    return x;
}
static void ufJoin(OWorld *w, uint32_t i, uint32_t j) {
    uint32_t a = i, b = j;
    for (;;) {
        a = ufFind(w, a);
        b = ufFind(w, b);
        if (a == b) return;
        uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        uint32_t old = CNT(hi, G_REGION);
        if (lo < old) CNT(hi, G_REGION) = lo;
        if (old == hi) return;
        a = old;
        b = lo;
    }
}
static inline int pinnedAt(OWorld *w, uint32_t i) { return ufLink(w, i) == i; }

/* ------------------------------------------------------------- particles */
typedef struct {
    v2 pos, vel;
    uint32_t word;
    float temp;
    uint32_t id;
    float affine[4];
} Particle;

static Particle loadP(const OWorld *w, const uint32_t *pb, uint32_t k) {
    Particle q;
    q.pos = V2(fbits(MOTION(pb, k, 0)), fbits(MOTION(pb, k, 1)));
    q.vel = V2(fbits(MOTION(pb, k, 2)), fbits(MOTION(pb, k, 3)));
    q.word = STATE(pb, k, 0);
    q.temp = fbits(STATE(pb, k, 1));
    q.id = STATE(pb, k, 2);
    for (int j = 0; j < 4; j++) q.affine[j] = fbits(AFFINE(pb, k, j));
    return q;
}
static void storeP(OWorld *w, uint32_t k, const Particle *q) {    /* :534-538 */
    uint32_t *pb = w->parts;
    MOTION(pb, k, 0) = ubits(q->pos.x);
    MOTION(pb, k, 1) = ubits(q->pos.y);
    MOTION(pb, k, 2) = ubits(q->vel.x);
    MOTION(pb, k, 3) = ubits(q->vel.y);
    STATE(pb, k, 0) = q->word;
    STATE(pb, k, 1) = ubits(q->temp);
    STATE(pb, k, 2) = q->id;
    STATE(pb, k, 3) = 0u;
    for (int j = 0; j < 4; j++) AFFINE(pb, k, j) = ubits(q->affine[j]);
}
static inline uint32_t pcount(const uint32_t *pb) { return pb[0]; }
static inline uint32_t pslots(const OWorld *w, const uint32_t *pb) { return pb[0] < w->cap ? pb[0] : w->cap; }

/* ------------------------------------------------------------------ sync */
static inline uint32_t claimOf(uint32_t i) { return DEST_CELL | (DEST_SPAN - i); }
static inline uint32_t claimedCell(uint32_t d) { return DEST_SPAN - (d & DEST_SPAN); }
static inline int isClaim(uint32_t d) { return d != DEST_KEEP && (d & DEST_ASIDE) == 0u && (d & DEST_CELL) != 0u; }

static void markBlockOf(OWorld *w, v2 pos) {                 /* :597-602 */
    i2 b = idiv8(cellOf(w, pos));
    if (inBlocks(w, b)) at_or(&w->blocks[blockIndex(w, b)], MARK_LIQUID);
}

/* a cell of a held body that was wet in the last solve (:583-595) */
static inline int wetRegionAt(const OWorld *w, const uint32_t *v) {
    uint32_t b = rigidBodyOf(w, v);
    return b != NONE && (bodyWord(w, b, BW_FLAGS) & (BF_WET | BF_HELD)) == (BF_WET | BF_HELD);
}
static void k_sync(OWorld *w, int claims) {                  /* :608-659 */
    PAR_FOR
    for (uint32_t by = 0; by < w->bh; by++)
        for (uint32_t bx = 0; bx < w->bw; bx++) {
            int liquidHere = 0, regionHere = 0;
            for (int ly = 0; ly < BLOCK; ly++)
                for (int lx = 0; lx < BLOCK; lx++) {
                    i2 c = I2((int32_t)bx * BLOCK + lx, (int32_t)by * BLOCK + ly);
                    if (!inWorld(w, c)) continue;
                    uint32_t i = cellIndex(w, c);
                    if (i == 0u) {
                        w->parts[0] = 0u;
                        w->blocks[COUNT_SLOT(COUNT_ACTIVE)] = 0u;
                        w->blocks[COUNT_SLOT(COUNT_RETIRED)] = 0u;
                        w->blocks[COUNT_SLOT(COUNT_VISCOUS)] = 0u;
                    }
                    const uint32_t *v = &w->cells[4ull * i];
                    if (kindOf(w, wType(v[0])) == KIND_LIQUID) {
                        uint32_t home = v[3] >> 8;
                        if (claims && wHas(v[0], FL_DRAWN) && home < w->C) {
                            uint32_t cl = claimOf(i);
                            at_max(&CNT(home, G_DEST), cl);
                        }
                        liquidHere = 1;
                    } else if (wetRegionAt(w, v)) {
                        regionHere = 1;
                    }
                }
            if (liquidHere) w->blocks[by * w->bw + bx] |= MARK_LIQUID;
            if (regionHere) w->blocks[by * w->bw + bx] |= MARK_REGION;
        }
}

static v2 pairWay(uint32_t a, uint32_t b) {                 /* :703-707 */
    uint32_t h = pcg((a < b ? a : b) ^ pcg(a < b ? b : a));
    float angle = (float)h * (float)(6.2831853 / 4294967296.0);
    float s = a < b ? 1.0f : -1.0f;
    return V2(cosf(angle) * s, sinf(angle) * s);
}

static v2 apart(OWorld *w, uint32_t k, uint32_t h, v2 pos, uint32_t id) {   /* :713-734 */
    uint32_t start = CNT(h, G_START), n = CNT(h, G_SORT);
    if (n < 2u) return pos;
    int32_t px = 0, py = 0;
    for (uint32_t m = start; m < start + n; m++) {
        if (m == k) continue;
        v2 d = V2(pos.x - fbits(MOTION(w->sorted, m, 0)), pos.y - fbits(MOTION(w->sorted, m, 1)));
        float l = wlength(d);
        if (l < APART) {
            v2 way = l == 0.0f ? pairWay(id, STATE(w->sorted, m, 2)) : V2(d.x / l, d.y / l);
            float s = MUTANT(16) ? 0.0f : 0.5f * (APART - l) * APART_FIX;
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

static uint32_t rankById(OWorld *w, uint32_t h, uint32_t k, uint32_t id, uint32_t word, uint32_t drawn) {  /* :668-678 */
    uint32_t start = CNT(h, G_START);
    uint32_t mine = wType(word) == drawn ? 1u : 0u, r = 0;
    for (uint32_t m = start; m < start + CNT(h, G_SORT); m++) {
        uint32_t ox = STATE(w->sorted, m, 0), oz = STATE(w->sorted, m, 2);
        uint32_t theirs = wType(ox) == drawn ? 1u : 0u;
        if (MUTANT(17)) r += (theirs < mine || (theirs == mine && m < k)) ? 1u : 0u;
        else r += (theirs < mine || (theirs == mine && (oz < id || (oz == id && m < k)))) ? 1u : 0u;
    }
    return r;
}

static uint32_t sharedLife(const OWorld *w, uint32_t life, uint32_t was, uint32_t now, uint32_t id) {  /* :4149-4161 */
    if (now == was) return life;
    float l = (float)life;
    if (now < was) l = l * (float)now / (float)was;
    else l = 255.0f - (255.0f - l) * (255.0f - (float)now) / (255.0f - (float)was);
    Rng r;
    seedRng(&r, id, w->P->frame, w->P->seed ^ 0x5a17u);
    uint32_t v = wu32(floorf(l + rnd(&r)));
    return v < 255u ? v : 255u;
}

static void k_syncParticles(OWorld *w) {                    /* :767-831 */
    uint32_t n = pslots(w, w->sorted);
    PAR_FOR
    for (uint32_t k = 0; k < n; k++) {
        Particle q = loadP(w, w->sorted, k);
        int keep = 0;
        uint32_t h = cellIndex(w, cellOf(w, q.pos));
        q.pos = apart(w, k, h, q.pos, q.id);
        uint32_t d = CNT(h, G_DEST);
        if (d == DEST_KEEP) {
            keep = 1;
            markBlockOf(w, q.pos);
        } else if ((d & DEST_ASIDE) != 0u) {
            uint32_t to = d & ~DEST_ASIDE;                        /* not into another body's cup (:785-794) */
            if (cupOf(w, to) == NONE)
                q.pos = V2((float)(to % (uint32_t)w->W) + (q.pos.x - floorf(q.pos.x)),
                           (float)(to / (uint32_t)w->W) + (q.pos.y - floorf(q.pos.y)));
            keep = 1;
            markBlockOf(w, q.pos);
        } else if (isClaim(d)) {
            uint32_t to = claimedCell(d);
            const uint32_t *v = &w->cells[4ull * to];
            uint32_t cnt = v[3] & LIQUID_COUNT_MAX;
            if (cnt == LIQUID_COUNT_MAX || CNT(h, G_SORT) <= cnt || rankById(w, h, k, q.id, q.word, wType(v[0])) < cnt) {
                if (to != h)
                    q.pos = V2((float)(to % (uint32_t)w->W) + (q.pos.x - floorf(q.pos.x)),
                               (float)(to / (uint32_t)w->W) + (q.pos.y - floorf(q.pos.y)));
                q.temp = fbits(v[1]);
                uint32_t drawn = CNT(h, G_TYPE), pt = wType(q.word);
                if (holdsSolute(pt) && pt == (drawn & 0xffu) && wType(v[0]) == pt
                    && (wAux(q.word) == ((drawn >> 8) & 0xffu) || wAux(q.word) == 0u)) {
                    uint32_t life = sharedLife(w, wAux(q.word) == 0u ? 0u : wLife(q.word), (drawn >> 16) & 0xffu, wLife(v[0]), q.id);
                    uint32_t ww = wWithout(v[0], FL_DRAWN);
                    uint32_t solute = wAux(ww) != 0u ? wAux(ww) : (drawn >> 8) & 0xffu;
                    q.word = packWord(pt, life > 0u ? solute : 0u, life, wShade(ww), wFlags(ww));
                } else if ((wType(q.word) | (wAux(q.word) << 8)) == (drawn & 0xffffu)) {
                    q.word = wWithout(v[0], FL_DRAWN);
                }
                keep = 1;
            }
        }
        if (keep) {
            uint32_t m = at_take(&w->parts[0], 1u);
            if (m < w->cap) storeP(w, m, &q);
        }
    }
}

static uint32_t seedId(const OWorld *w, i2 c, uint32_t j) {  /* :835-841 */
    uint32_t h = (((uint32_t)c.y * (uint32_t)w->W + (uint32_t)c.x) * 0x9e3779b9u) ^ (w->P->frame * 0x85ebca6bu) ^ (j * 0xc2b2ae35u);
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    return h;
}
static v2 seedPos(Rng *r, i2 c, uint32_t j, uint32_t n) {   /* :845-855 */
    if (n == 1u) return V2((float)c.x + 0.5f, (float)c.y + 0.5f);
    float rx = rnd(r), ry = rnd(r);
    v2 q = V2(rx - 0.5f, ry - 0.5f);
    if (n <= 4u) {
        static const float quarter[4][2] = {{0, 0}, {1, 1}, {1, 0}, {0, 1}};
        return V2((float)c.x + 0.25f + 0.5f * quarter[j][0] + 0.3f * q.x, (float)c.y + 0.25f + 0.5f * quarter[j][1] + 0.3f * q.y);
    }
    return V2((float)c.x + 0.5f + 0.9f * q.x, (float)c.y + 0.5f + 0.9f * q.y);
}
static void k_syncSeed(OWorld *w) {                          /* :863-890 */
    PAR_FOR
    for (uint32_t by = 0; by < w->bh; by++)
        for (uint32_t bx = 0; bx < w->bw; bx++)
            for (int ly = 0; ly < BLOCK; ly++)
                for (int lx = 0; lx < BLOCK; lx++) {
                    i2 c = I2((int32_t)bx * BLOCK + lx, (int32_t)by * BLOCK + ly);
                    if (!inWorld(w, c)) continue;
                    uint32_t i = cellIndex(w, c), n = 0;
                    const uint32_t *v = &w->cells[4ull * i];
                    Rng r = {0};
                    if (kindOf(w, wType(v[0])) != KIND_LIQUID) continue;
                    uint32_t home = v[3] >> 8, want = v[3] & LIQUID_COUNT_MAX;
                    if (!wHas(v[0], FL_DRAWN)) n = LIQUID_PER_CELL;
                    else if (home >= w->C || CNT(home, G_DEST) != claimOf(i)) n = want > 1u ? want : 1u;
                    else if (want > CNT(home, G_SORT) && want < LIQUID_COUNT_MAX) n = want - CNT(home, G_SORT);
                    seedRng(&r, i, w->P->frame, w->P->seed ^ 0x5eedu);
                    uint32_t m = at_take(&w->parts[0], n);
                    for (uint32_t j = 0; j < n; j++) {
                        Particle q;
                        q.pos = seedPos(&r, c, j, n);
                        q.vel = unpack2x16(v[2]);
                        q.word = wWithout(v[0], FL_DRAWN);
                        q.temp = fbits(v[1]);
                        q.id = seedId(w, c, j);
                        q.affine[0] = q.affine[1] = q.affine[2] = q.affine[3] = 0.0f;
                        if (m + j < w->cap) storeP(w, m + j, &q);
                    }
                }
}

/* ------------------------------------------------------------------ plan */
static void k_compact(OWorld *w) {                           /* :896-921 */
    for (uint32_t n = 0; n < w->B; n++) {
        i2 b = I2((int32_t)(n % w->bw), (int32_t)(n / w->bw));
        uint32_t marks = 0;
        const int reach = MUTANT(20) ? 0 : 1;
        for (int y = -reach; y <= reach; y++)
            for (int x = -reach; x <= reach; x++) {
                i2 o = I2(b.x + x, b.y + y);
                if (inBlocks(w, o)) marks |= w->blocks[blockIndex(w, o)];
            }
        int near = (marks & MARK_LIQUID) != 0u, liquid = (marks & (MARK_LIQUID | MARK_REGION)) != 0u;
        int was = (w->blocks[STATE_SLOT(n)] & ACTIVE) != 0u;
        w->blocks[STATE_SLOT(n)] = (liquid ? ACTIVE : 0u) | (near ? NEAR_LIQUID : 0u);
        if (liquid) w->blocks[ACTIVE_SLOT(w->blocks[COUNT_SLOT(COUNT_ACTIVE)]++)] = n;
        else if (was) w->blocks[RETIRED_SLOT(w->blocks[COUNT_SLOT(COUNT_RETIRED)]++)] = n;
    }
}

static void clearCell(OWorld *w, uint32_t i) {
    for (uint32_t g = 0; g < LIQUID_CELL_WORDS; g++) CNT(i, g) = 0u;
}

static void k_retire(OWorld *w) {                            /* :924-951 */
    PAR_FOR
    for (uint32_t wid = 0; wid < w->blocks[COUNT_SLOT(COUNT_RETIRED)]; wid++) {
        uint32_t n = w->blocks[RETIRED_SLOT(wid)];
        for (int ly = 0; ly < BLOCK; ly++)
            for (int lx = 0; lx < BLOCK; lx++) {
                i2 c = I2((int32_t)(n % w->bw) * BLOCK + lx, (int32_t)(n / w->bw) * BLOCK + ly);
                if (!inWorld(w, c)) continue;
                uint32_t i = cellIndex(w, c);
                for (int j = 0; j < 4; j++) { FLD(2u * i, j) = 0.0f; FLD(2u * i + 1u, j) = 0.0f; }
                clearCell(w, i);
                Face3 fs[4];
                facesOf(w, c, fs);
                for (int j = 0; j < 4; j++) {
                    if (fs[j].isU < 0) continue;
                    uint32_t f = faceIndex(w, I2(fs[j].x, fs[j].y), fs[j].isU == 1);
                    for (int q = 0; q < 4; q++) FLD(faceSlot(w, f), q) = 0.0f;
                    FLD(viscLastSlot(w, f), f % 4u) = 0.0f;
                    SUM(f, 0) = 0u;
                    SUM(f, 1) = 0u;
                }
            }
        w->blocks[n] = 0u;
    }
}

static void k_clear(OWorld *w) {                             /* :962-985 */
    PAR_FOR
    for (uint32_t wid = 0; wid < activeCount(w); wid++) {
        i2 b = activeBlock(w, wid);
        w->blocks[blockIndex(w, b)] = 0u;
        for (int ly = 0; ly < BLOCK; ly++)
            for (int lx = 0; lx < BLOCK; lx++) {
                i2 c = I2(b.x * BLOCK + lx, b.y * BLOCK + ly);
                if (!inWorld(w, c)) continue;
                uint32_t i = cellIndex(w, c);
                clearCell(w, i);
                CNT(i, G_REGION) = i;
                Face3 fs[4];
                facesOf(w, c, fs);
                for (int j = 0; j < 4; j++) {
                    if (fs[j].isU < 0) continue;
                    uint32_t f = faceIndex(w, I2(fs[j].x, fs[j].y), fs[j].isU == 1);
                    SUM(f, 0) = 0u;
                    SUM(f, 1) = 0u;
                }
            }
    }
}

/* ---------------------------------------------------------------- gather */
static void splatAxis(OWorld *w, v2 s, float vel, float gx, float gy, int isU) {   /* :1000-1015 */
    float fx0 = floorf(s.x), fy0 = floorf(s.y);
    i2 b = I2(wi32(fx0), wi32(fy0));
    v2 f = V2(s.x - fx0, s.y - fy0);
    for (int j = 0; j < 4; j++) {
        int ox = j & 1, oy = j >> 1;
        float wgt = (ox == 1 ? f.x : 1.0f - f.x) * (oy == 1 ? f.y : 1.0f - f.y);
        i2 at = I2(b.x + ox, b.y + oy);
        if (wgt <= 0.0f || !faceInWorld(w, at, isU)) continue;
        uint32_t fi = faceIndex(w, at, isU);
        float there = vel + (gx * ((float)at.x - s.x) + gy * ((float)at.y - s.y));
        at_add(&SUM(fi, 0), (uint32_t)wi32(wround(there * wgt * FIX)));
        at_add(&SUM(fi, 1), (uint32_t)wi32(wround(wgt * FIX)));
    }
}
static void splatDensity(OWorld *w, v2 pos, float room) {   /* :1069-1087 */
    v2 s = V2(pos.x - 0.5f, pos.y - 0.5f);
    float fx0 = floorf(s.x), fy0 = floorf(s.y);
    i2 b = I2(wi32(fx0), wi32(fy0));
    v2 f = V2(s.x - fx0, s.y - fy0);
    for (int j = 0; j < 4; j++) {
        int ox = j & 1, oy = j >> 1;
        float wgt = (ox == 1 ? f.x : 1.0f - f.x) * (oy == 1 ? f.y : 1.0f - f.y);
        i2 at = I2(b.x + ox, b.y + oy);
        if (walls(w, at)) {
            i2 x = I2(b.x + 1 - ox, at.y), y = I2(at.x, b.y + 1 - oy), xy = I2(b.x + 1 - ox, b.y + 1 - oy);
            at = !walls(w, x) ? x : (!walls(w, y) ? y : (!walls(w, xy) ? xy : I2(-1, -1)));
        }
        if (inWorld(w, at) && wgt > 0.0f) at_add(&CNT(cellIndex(w, at), G_DENSITY), wu32(wround(wgt * room * FIX)));
    }
}
static void k_gather(OWorld *w) {                            /* :1017-1039 */
    uint32_t n = pslots(w, w->parts);
    PAR_FOR
    for (uint32_t k = 0; k < n; k++) {
        Particle q = loadP(w, w->parts, k);
        /* the air's pressure on a particle beside gas (:1024-1032): its
         * velocity pushed, and kept */
        if (w->ctl->air) {
            i2 c = I2(wi32(floorf(q.pos.x)), wi32(floorf(q.pos.y)));
            int gas = 0;
            for (int d = 0; d < 4; d++) {
                i2 n = beside(c, d);
                gas |= inWorld(w, n) && kindOf(w, wType(w->cells[4ull * cellIndex(w, n)])) == KIND_GAS;
            }
            if (gas) {
                v2 push = pressurePush(w, airGrad(w, q.pos), WATER_RHO);
                q.vel.x += push.x;
                q.vel.y += push.y;
                MOTION(w->parts, k, 2) = ubits(q.vel.x);
                MOTION(w->parts, k, 3) = ubits(q.vel.y);
            }
        }
        splatAxis(w, V2(q.pos.x - 0.0f, q.pos.y - 0.5f), q.vel.x, q.affine[0], q.affine[1], 1);
        splatAxis(w, V2(q.pos.x - 0.5f, q.pos.y - 0.0f), q.vel.y, q.affine[2], q.affine[3], 0);
        uint32_t home = cellIndex(w, cellOf(w, q.pos));
        at_add(&CNT(home, G_COUNT), 1u);
        at_add(&CNT(home, G_MASS), wu32(wround(liquidDensity(q.word, q.temp) * FIX)));
        splatDensity(w, q.pos, liquidVolume(q.word, q.temp));
    }
}

/* ------------------------------------------------------------------ seal */
static int mayBeSealed(OWorld *w, i2 c) {                   /* :1513-1519 */
    if (countAt(w, c) != 0) return 0;
    uint32_t i = cellIndex(w, c);
    return FLD(2u * i, 2) > 0.0f || FLD(2u * i + 1u, 2) != 0.0f;
}
static int wasOpen(OWorld *w, i2 q) {                       /* :1503-1508 */
    if (!inWorld(w, q) || countAt(w, q) < 0) return 0;
    return (wu32(FLD(2u * cellIndex(w, q) + 1u, 3)) & stepBit(w, 1u)) == 0u;
}
#define FOR_ACTIVE_CELLS_SERIAL(w, c)                                            \
    for (uint32_t wid_ = 0; wid_ < activeCount(w); wid_++)                       \
        for (int ly_ = 0; ly_ < BLOCK; ly_++)                                    \
            for (int lx_ = 0; lx_ < BLOCK; lx_++)                                \
                for (i2 c = I2(activeBlock(w, wid_).x * BLOCK + lx_, activeBlock(w, wid_).y * BLOCK + ly_), \
                     *once_ = &c; once_; once_ = NULL)
#define FOR_ACTIVE_CELLS(w, c) PAR_FOR FOR_ACTIVE_CELLS_SERIAL(w, c)

static void k_sealJoin(OWorld *w) {                          /* :1531-1543 */
    FOR_ACTIVE_CELLS_SERIAL(w, c) {
        if (!inWorld(w, c) || !mayBeSealed(w, c)) continue;
        for (int d = 0; d < 2; d++) {
            i2 n = d == 0 ? I2(c.x + 1, c.y) : I2(c.x, c.y + 1);
            if (mayBeSealed(w, n)) ufJoin(w, cellIndex(w, c), cellIndex(w, n));
        }
    }
}
static void k_sealAnchor(OWorld *w) {                        /* :1545-1554 */
    FOR_ACTIVE_CELLS_SERIAL(w, c) {
        if (!inWorld(w, c) || !mayBeSealed(w, c)) continue;
        if (wasOpen(w, beside(c, 0)) || wasOpen(w, beside(c, 1)) || wasOpen(w, beside(c, 2)) || wasOpen(w, beside(c, 3)))
            CNT(ufFind(w, cellIndex(w, c)), G_REGION) |= REGION_OPEN;
    }
}

static float columnWeight(OWorld *w, i2 p, float density) {  /* :1566-1575 */
    int over = 0;
    i2 q = I2(p.x, p.y - 1);
    while (over < COLUMN_MOST && countAt(w, q) > 0) {
        over += 1;
        q.y -= 1;
    }
    return density * w->P->earthG * ((float)over + 0.5f);
}

/* surface tension (:1698-1806): waterTension, the liquid's outline, and the
 * surface pass, σκ on the open cells beside the liquid */
#define CAPILLARY_LENGTH 0.0027f                             /* js/shared.js:285-286 */
#define SURFACE_MIN_LENGTH 0.3f
#define MAX_CURVATURE 1.0f
#define CURVATURE_GAIN 2.0f
#define CONTACT 2.0f
#define SURFACE_DAMP 1.0f
static float waterTension(const OWorld *w) {                 /* :1736-1739 */
    float l = CAPILLARY_LENGTH / w->P->dx;
    return l >= SURFACE_MIN_LENGTH ? WATER_WEIGHT * l * l : 0.0f;
}
static float outline(const OWorld *w, i2 c) {                 /* :1744-1754 */
    if (!inWorld(w, c)) return 0.5f;
    uint32_t i = cellIndex(w, c);
    if (isObstacleCell(w, c, w->cells[4ull * i])) return 0.5f;
    return CNT(i, G_COUNT) > 0u ? 1.0f : 0.0f;
}
static void k_surface(OWorld *w) {                           /* :1760-1807 */
    FOR_ACTIVE_CELLS(w, c) {
        if (!inWorld(w, c)) continue;
        uint32_t i = cellIndex(w, c);
        if (FLD(2u * i, 2) != 0.0f) continue;
        float tension = 0.0f, contact = 0.0f;
        for (int d = 0; d < 4; d++) {
            i2 q = beside(c, d);
            if (!inWorld(w, q)) continue;
            uint32_t j = cellIndex(w, q);
            float z = FLD(2u * j, 2);
            if (z > 0.0f) tension = fmaxf(tension, w->props[wType(w->cells[4ull * j])].tension);
            else if (z < 0.0f) contact += 2.0f * w->props[wType(w->cells[4ull * j])].wetting - 1.0f;
        }
        if (tension <= 0.0f) continue;
        float gx = 0.0f, gy = 0.0f, hx = 0.0f, hy = 0.0f, hz = 0.0f;
        for (int y = -2; y <= 2; y++)
            for (int x = -2; x <= 2; x++) {
                float kx = (float)x, ky = (float)y;
                float wt = wexpf(-0.5f * (kx * kx + ky * ky)) * outline(w, I2(c.x + x, c.y + y));
                gx += kx * wt;
                gy += ky * wt;
                hx += (kx * kx - 1.0f) * wt;
                hy += (ky * ky - 1.0f) * wt;
                hz += (kx * ky) * wt;
            }
        float g2 = gx * gx + gy * gy, kappa = 0.0f;
        if (g2 >= 1e-6f) kappa = -(hx * gy * gy - 2.0f * gx * gy * hz + hy * gx * gx) / (g2 * sqrtf(g2));
        if (MUTANT(34)) kappa = -kappa;
        float curved = wclamp(kappa * CURVATURE_GAIN, -MAX_CURVATURE, MAX_CURVATURE);
        FLD(2u * i, 0) = waterTension(w) * tension * (curved - CONTACT * contact);
    }
}
static void k_classify(OWorld *w) {                          /* :1578-1696 */
    PAR_FOR
    for (uint32_t wid = 0; wid < activeCount(w); wid++) {
        i2 b = activeBlock(w, wid);
        uint32_t blockBits = 0;
        for (int ly = 0; ly < BLOCK; ly++)
            for (int lx = 0; lx < BLOCK; lx++) {
                i2 p = I2(b.x * BLOCK + lx, b.y * BLOCK + ly);
                if (!inWorld(w, p)) continue;
                uint32_t i = cellIndex(w, p);
                int32_t n = countAt(w, p);
                float c[4] = {0, 0, 0, 0};
                float was[4], wasDrift[4];
                for (int j = 0; j < 4; j++) { was[j] = FLD(2u * i, j); wasDrift[j] = FLD(2u * i + 1u, j); }
                int sealed = 0;
                if (n == 0 && mayBeSealed(w, p) && !MUTANT(15)) sealed = (ufLink(w, ufFind(w, i)) & REGION_OPEN) == 0u;
                uint32_t rigid = NONE, held = NONE;
                uint32_t seam = n < 0 ? NONE : seamOf(w, i);
                if (seam != NONE) {
                    /* a gap the stamp opened in a body's wall: the body's (:1603-1612) */
                    c[2] = -1.0f;
                    rigid = seam;
                    if (isHeld(w, seam)) {
                        held = seam;
                        rigid = NONE;
                    }
                    sealed = 0;
                } else if (n < 0) {
                    c[2] = -1.0f;
                    rigid = rigidBodyOf(w, &w->cells[4ull * i]);
                    /* powder in its cup is its cargo, moving with it (:1616-1619) */
                    if (rigid == NONE && kindOf(w, wType(w->cells[4ull * i])) == KIND_POWDER) rigid = cupOf(w, i);
                    if (rigid != NONE && isHeld(w, rigid)) {
                        held = rigid;
                        rigid = NONE;
                    }
                } else if (n > 0) {
                    v2 li = liquidIn(w, i);
                    c[2] = li.x;
                    c[3] = li.y;
                    /* liquid at the surface, where surface tension acts, takes on
                     * its viscosity (SURFACE_DAMP, :1626-1633) */
                    float sigma = waterTension(w);
                    if (sigma > 0.0f && (countAt(w, beside(p, 0)) == 0 || countAt(w, beside(p, 1)) == 0
                                         || countAt(w, beside(p, 2)) == 0 || countAt(w, beside(p, 3)) == 0))
                        c[3] = fmaxf(c[3], sigma * w->props[wType(w->cells[4ull * i])].tension * SURFACE_DAMP);
                } else {
                    float fluid = 0.0f, nwalls = 0.0f;
                    v2 mean = V2(0.0f, 0.0f);
                    for (int d = 0; d < 4; d++) {
                        i2 q = beside(p, d);
                        int32_t m = countAt(w, q);
                        if (m < 0) nwalls += 1.0f;
                        else if (m > 0) {
                            fluid += 1.0f;
                            v2 li = liquidIn(w, cellIndex(w, q));
                            mean.x += li.x;
                            mean.y += li.y;
                        }
                    }
                    if (fluid >= (MUTANT(14) ? 1.0f : 2.0f) && fluid + nwalls >= 3.0f) {
                        c[2] = mean.x / fluid;
                        c[3] = mean.y / fluid;
                    } else if (sealed) {
                        c[2] = fluid > 0.0f ? mean.x / fluid : 1.0f;
                        c[3] = fluid > 0.0f ? mean.y / fluid : 0.0f;
                    }
                }
                uint32_t mine = stepBit(w, 0u);
                uint32_t liquid = (c[2] > 0.0f || rigid != NONE || held != NONE) ? mine : 0u;
                float drift[4] = {0.0f, 0.0f, sealed ? 1.0f : 0.0f,
                                  (float)((wu32(wasDrift[3]) & ~(mine | fallingBit(w, 0u))) | liquid)};
                if (rigid != NONE) drift[2] = -1.0f - (float)rigid;
                else if (held != NONE) drift[2] = -1.0f - (float)(held + MAX_BODIES);
                if (c[2] > 0.0f) {
                    if (was[2] > 0.0f) {
                        c[0] = was[0];
                        drift[0] = wasDrift[0] * DRIFT_KEEP;
                    } else {
                        c[0] = columnWeight(w, p, c[2]);
                    }
                    blockBits |= c[3] > 0.0f ? (HAS_FLUID | HAS_VISCOUS) : HAS_FLUID;
                }
                for (int j = 0; j < 4; j++) { FLD(2u * i, j) = c[j]; FLD(2u * i + 1u, j) = drift[j]; }
            }
        w->blocks[STATE_SLOT(blockIndex(w, b))] |= blockBits;
        if (blockBits & HAS_VISCOUS) at_or(&w->blocks[COUNT_SLOT(COUNT_VISCOUS)], 1u);
    }
}

/* ----------------------------------------------------------------- faces */
static float obstacleVel(OWorld *w, i2 c, int axis) {       /* :1820-1834 */
    if (!inWorld(w, c)) return 0.0f;
    uint32_t held = occupantAt(w, cellIndex(w, c));
    if (isOccupant(held)) {
        v2 u = occupantVel(held);
        return axis == 0 ? u.x : u.y;
    }
    const uint32_t *v = &w->cells[4ull * cellIndex(w, c)];
    uint32_t k = kindOf(w, wType(v[0]));
    if (k == KIND_SOLID || k == KIND_GAS) return 0.0f;
    v2 u = unpack2x16(v[2]);
    return (axis == 0 ? u.x : u.y) * POWDER_PUSH;
}
static float faceViscosity(const float ga[4], const float gb[4]) {   /* :3279-3286 */
    if (ga[2] < 0.0f || gb[2] < 0.0f) return 0.0f;
    float fa = ga[2] > 0.0f ? 1.0f : 0.0f, fb = gb[2] > 0.0f ? 1.0f : 0.0f;
    return (fa * ga[3] + fb * gb[3]) / fmaxf(fa + fb, 1.0f);
}
/* wasLiquid (:3315-3317): cell c liquid last step too */
static inline int wasLiquid(const OWorld *w, i2 c) {
    return (wu32(FLD(2u * cellIndex(w, c) + 1u, 3)) & stepBit(w, 1u)) != 0u;
}
/* lastPush (:3305-3310): last step's pressure push across the face */
static float lastPush(const OWorld *w, i2 a, i2 at, const float ga[4], const float gb[4]) {
    if ((ga[2] > 0.0f && !wasLiquid(w, a)) || (gb[2] > 0.0f && !wasLiquid(w, at))) return 0.0f;
    return fineBeta(ga[2], gb[2], densityAt(w, a), densityAt(w, at)) * (gb[0] - ga[0]);
}
/* a face of a free body's cell: the body's velocity there going into the
 * solve, the mean between two bodies' cells (:1841-1858) */
static void setBodyFace(OWorld *w, uint32_t f, i2 at, int isU, const float ga[4], const float gb[4], uint32_t ra, uint32_t rb,
                        int drag) {
    int axis = isU ? 0 : 1;
    v2 x = faceMiddle(at, isU);
    float now, u;
    if (ra != NONE && rb != NONE) {
        v2 ma = bodyMoving(w, ra, x, drag), mb = bodyMoving(w, rb, x, drag), va = rigidVel(w, ra, x), vb = rigidVel(w, rb, x);
        now = 0.5f * ((axis ? ma.y : ma.x) + (axis ? mb.y : mb.x));
        u = 0.5f * ((axis ? va.y : va.x) + (axis ? vb.y : vb.x));
    } else {
        uint32_t b = ra != NONE ? ra : rb;
        v2 m = bodyMoving(w, b, x, drag), v = rigidVel(w, b, x);
        now = axis ? m.y : m.x;
        u = axis ? v.y : v.x;
    }
    const float *other = ra != NONE ? gb : ga;
    float nu = (ra == NONE || rb == NONE) && other[2] > 0.0f ? other[3] : 0.0f;
    FLD(faceSlot(w, f), 0) = now;
    FLD(faceSlot(w, f), 1) = u;
    FLD(faceSlot(w, f), 2) = 1.0f;
    FLD(faceSlot(w, f), 3) = nu;
    FLD(2ull * w->C + w->F + f / 4u, f % 4u) = now;          /* setViscRhs */
}
static void setFace(OWorld *w, i2 at, int isU) {           /* :1936-1980 */
    uint32_t f = faceIndex(w, at, isU);
    i2 a = isU ? I2(at.x - 1, at.y) : I2(at.x, at.y - 1);
    int axis = isU ? 0 : 1;
    float ga[4], gb[4];
    gridAt4(w, a, ga);
    gridAt4(w, at, gb);
    int32_t weight = (int32_t)SUM(f, 1);
    uint32_t ra = rigidAt(w, a), rb = rigidAt(w, at);
    if (ra != NONE || rb != NONE) {
        setBodyFace(w, f, at, isU, ga, gb, ra, rb, 0);
        return;
    }
    float u = 0.0f, held = 0.0f;
    if (ga[2] < 0.0f && gb[2] < 0.0f) {
        held = -1.0f;
    } else if (ga[2] < 0.0f || gb[2] < 0.0f) {
        u = obstacleVel(w, ga[2] < 0.0f ? a : at, axis);
        held = 1.0f;
    } else if (weight > 0) {
        u = (float)(int32_t)SUM(f, 0) / (float)weight;
        held = 1.0f;
    }
    float now = u;
    if (held > 0.0f && w->ctl->air) now += surfaceDrag(w, a, at, isU, ga, gb, u);   /* the wind's drag */
    if (held > 0.0f && ga[2] >= 0.0f && gb[2] >= 0.0f) {   /* the pull as it is at the face (:1968-1972) */
        v2 pull = pullAt(w, isU ? V2((float)at.x, (float)at.y + 0.5f) : V2((float)at.x + 0.5f, (float)at.y));
        now += (isU != MUTANT(21)) ? pull.x : pull.y;
    }
    float start = now;
    if (faceViscosity(ga, gb) > 0.0f) {                      /* isViscousFace: last step's push off (:1975-1978) */
        now -= lastPush(w, a, at, ga, gb);
        start = now + FLD(viscLastSlot(w, f), f % 4u);
    }
    FLD(faceSlot(w, f), 0) = start;
    FLD(faceSlot(w, f), 1) = u;
    FLD(faceSlot(w, f), 2) = held;
    FLD(faceSlot(w, f), 3) = faceViscosity(ga, gb);
    FLD(2ull * w->C + w->F + f / 4u, f % 4u) = now;          /* setViscRhs */
}
static void k_setFaces(OWorld *w) {                          /* :1982-1994 */
    FOR_ACTIVE_CELLS(w, c) {
        if (!inWorld(w, c)) continue;
        Face3 fs[4];
        facesOf(w, c, fs);
        for (int j = 0; j < 4; j++)
            if (fs[j].isU >= 0) setFace(w, I2(fs[j].x, fs[j].y), fs[j].isU == 1);
    }
}
/* viscousFace (:3347-3403): a red-black Gauss-Seidel step of implicit
 * viscosity on the face at `at` if it is of `color`; a body's face with
 * liquid beyond it too, the body's mass on its side (heavy) */
#define VISC_OMEGA 1.6f
#define VISCOSITY_ITERATIONS 8                               /* js/sim.js:71 */
static void viscousFace(OWorld *w, i2 at, int isU, int color) {
    if (((at.x + at.y) & 1) != color) return;
    i2 a = isU ? I2(at.x - 1, at.y) : I2(at.x, at.y - 1);
    float ga[4], gb[4];
    gridAt4(w, a, ga);
    gridAt4(w, at, gb);
    uint32_t ra = rigidAt(w, a), rb = rigidAt(w, at);
    int surface = (ra != NONE) != (rb != NONE) && (ra != NONE ? gb[2] : ga[2]) > 0.0f;
    if (!surface && (ga[2] < 0.0f || gb[2] < 0.0f || ra != NONE || rb != NONE)) return;
    uint32_t f = faceIndex(w, at, isU);
    uint64_t fs = faceSlot(w, f);
    float nu = FLD(fs, 3);
    if (nu <= 0.0f) return;
    float heavy = 1.0f;
    if (surface) {
        float liquid = ra != NONE ? gb[2] : ga[2];
        heavy = 2.0f * liquid / (liquid + bodyDensity(w, ra != NONE ? ra : rb));
    }
    float through[2] = {ga[2] > 0.0f ? ga[3] : 0.0f, gb[2] > 0.0f ? gb[3] : 0.0f};
    float sum = 0.0f, n = 0.0f;
    for (int d = 0; d < 4; d++) {
        i2 o = beside(at, d);
        int across = isU == (d >= 2);
        float other[4] = {0.0f, 0.0f, -1.0f, 0.0f};
        if (faceInWorld(w, o, isU)) {
            uint64_t os = faceSlot(w, faceIndex(w, o, isU));
            for (int j = 0; j < 4; j++) other[j] = FLD(os, j);
        }
        float k = (d & 1) ? through[1] : through[0];
        if (across) k = other[2] < 0.0f ? nu : 2.0f * nu * other[3] / fmaxf(nu + other[3], 1e-12f);
        if (other[2] < 0.0f) n += k * ((across && !MUTANT(33)) ? 2.0f : 1.0f);
        else if (other[2] > 0.0f) {
            sum += k * other[0];
            n += k;
        }
    }
    float rhs = FLD(2ull * w->C + w->F + f / 4u, f % 4u);   /* viscRhs */
    FLD(fs, 0) = wmix(FLD(fs, 0), (rhs + heavy * sum) / (1.0f + heavy * n), MUTANT(32) ? 1.0f : VISC_OMEGA);
}
/* viscosity (:3407-3447): over the blocks worked on with viscous liquid in
 * or beside them, each cell's faces: a sweep of one colour (0, 1) or
 * unpress (2): what the sweeps did kept for next step (viscLast) and last
 * step's push back on; on the faces they don't solve, viscLast cleared */
static void k_viscosity(OWorld *w, int color) {
    PAR_FOR
    for (uint32_t wid = 0; wid < activeCount(w); wid++) {
        i2 b = activeBlock(w, wid);
        if ((statesOf(w, I2(b.x - 1, b.y - 1), I2(b.x + 1, b.y + 1)) & HAS_VISCOUS) == 0u) continue;
        for (int li = 0; li < BLOCK * BLOCK; li++) {
            i2 c = I2(b.x * BLOCK + li % BLOCK, b.y * BLOCK + li / BLOCK);
            if (!inWorld(w, c)) continue;
            Face3 fs[4];
            facesOf(w, c, fs);
            for (int j = 0; j < 4; j++) {
                if (fs[j].isU < 0) continue;
                i2 at = I2(fs[j].x, fs[j].y);
                int isU = fs[j].isU == 1;
                if (color < 2) {
                    viscousFace(w, at, isU, color);
                    continue;
                }
                i2 a = isU ? I2(at.x - 1, at.y) : I2(at.x, at.y - 1);
                float ga[4], gb[4];
                gridAt4(w, a, ga);
                gridAt4(w, at, gb);
                uint32_t f = faceIndex(w, at, isU);
                if (rigidAt(w, a) == NONE && rigidAt(w, at) == NONE && faceViscosity(ga, gb) > 0.0f) {
                    uint64_t fs0 = faceSlot(w, f);
                    FLD(viscLastSlot(w, f), f % 4u) = FLD(fs0, 0) - FLD(2ull * w->C + w->F + f / 4u, f % 4u);
                    FLD(fs0, 0) += lastPush(w, a, at, ga, gb);
                } else if (FLD(viscLastSlot(w, f), f % 4u) != 0.0f) {
                    FLD(viscLastSlot(w, f), f % 4u) = 0.0f;
                }
            }
        }
    }
}
static void k_unpress(OWorld *w) { k_viscosity(w, 2); }   /* :3449-3454 */
static void k_divergence(OWorld *w) {                        /* :1996-2025 */
    w->blocks[COUNT_SLOT(COUNT_RESIDUAL)] = 0u;
    w->blocks[COUNT_SLOT(COUNT_CYCLES)] = 0u;
    FOR_ACTIVE_CELLS(w, c) {
        if (!inWorld(w, c)) continue;
        uint32_t i = cellIndex(w, c);
        if (FLD(2u * i, 2) <= 0.0f) continue;
        float out = FLD(faceSlot(w, uFace(w, c.x + 1, c.y)), 0) - FLD(faceSlot(w, uFace(w, c.x, c.y)), 0)
            + FLD(faceSlot(w, vFace(w, c.x, c.y + 1)), 0) - FLD(faceSlot(w, vFace(w, c.x, c.y)), 0);
        FLD(2u * i, 1) = out;
        float n = densityAt(w, c);
        float room = 1.0f;                                   /* posedRoom (:2035-2045) */
        for (int d = 0; d < 4; d++) {
            i2 q = beside(c, d);
            uint32_t b = rigidAt(w, q);
            if (b != NONE) room += 0.5f - bodyReach(w, b, I2(c.x - q.x, c.y - q.y));
        }
        room = wclamp(room, 0.1f, 1.9f);
        FLD(2u * i + 1u, 1) = FLD(2u * i + 1u, 2) < 0.0f ? 0.0f : (MUTANT(24) ? SPREAD : -SPREAD) * (n - room);
    }
}

/* ------------------------------------------------------------- multigrid */
static inline int converged(const OWorld *w) { return w->cycle > 0u && w->converged_flag; }
static inline int inLevel(const OWorld *w, uint32_t l, i2 c) {
    return c.x >= 0 && c.y >= 0 && c.x < (int32_t)w->lw[l] && c.y < (int32_t)w->lh[l];
}
static inline uint64_t levelSlot(const OWorld *w, uint32_t l, i2 c) {
    return w->lstart[l] + 2ull * ((uint64_t)c.y * w->lw[l] + (uint64_t)c.x);
}
static inline float rhoAt(const OWorld *w, uint32_t l, i2 c) {
    if (!inLevel(w, l, c)) return -1.0f;
    if (l == 0) return FLD(2u * cellIndex(w, c), 2);
    return FLD(levelSlot(w, l, c) + 1u, 0);
}
static inline v2 potAt(const OWorld *w, uint32_t l, i2 c) {
    if (!inLevel(w, l, c)) return V2(0.0f, 0.0f);
    if (l == 0) { uint32_t i = 2u * cellIndex(w, c); return V2(FLD(i, 0), FLD(i + 1u, 0)); }
    uint64_t at = levelSlot(w, l, c);
    return V2(FLD(at, 0), FLD(at, 1));
}
static inline v2 rhsAt(const OWorld *w, uint32_t l, i2 c) {
    if (l == 0) { uint32_t i = 2u * cellIndex(w, c); return V2(FLD(i, 1), FLD(i + 1u, 1)); }
    uint64_t at = levelSlot(w, l, c);
    return V2(FLD(at, 2), FLD(at, 3));
}
typedef struct { v2 sum; float diag; } Stencil;
static Stencil stencil(const OWorld *w, uint32_t l, i2 c, float rho) {   /* :2286-2302 */
    Stencil s = {{0.0f, 0.0f}, 0.0f};
    for (int d = 0; d < 4; d++) {
        i2 n = beside(c, d);
        float r = rhoAt(w, l, n);
        if (r < 0.0f) continue;
        float beta = faceBeta(rho, r);
        if (l == 0) beta = fineBeta(rho, r, densityAt(w, c), densityAt(w, n));
        v2 p = potAt(w, l, n);
        s.sum.x += beta * p.x;
        s.sum.y += beta * p.y;
        s.diag += beta;
    }
    return s;
}
static v2 residualAt(OWorld *w, uint32_t l, i2 c) {         /* :2306-2315 */
    float rho = rhoAt(w, l, c);
    if (rho <= 0.0f) return V2(0.0f, 0.0f);
    Stencil s = stencil(w, l, c, rho);
    float h2 = (float)(1u << (2u * l));
    v2 pin = (l == 0 && pinnedAt(w, cellIndex(w, c))) ? V2(PIN, 0.0f) : V2(0.0f, 0.0f);
    v2 pot = potAt(w, l, c), rhs = rhsAt(w, l, c);
    return V2(rhs.x - (s.sum.x - (s.diag + pin.x) * pot.x) / h2, rhs.y - (s.sum.y - (s.diag + pin.y) * pot.y) / h2);
}
static void smoothAt(OWorld *w, uint32_t l, i2 c, int color, float omega) {   /* :2318-2334 */
    if (!inLevel(w, l, c) || ((c.x + c.y) & 1) != color || activeCount(w) == 0u || converged(w)) return;
    uint64_t at = levelSlot(w, l, c);
    float rho = FLD(at + 1u, 0);
    if (rho <= 0.0f) return;
    Stencil s = stencil(w, l, c, rho);
    if (s.diag <= 0.0f) return;
    float h2 = MUTANT(5) ? 1.0f : (float)(1u << (2u * l));
    float ex = FLD(at, 0), ey = FLD(at, 1), ez = FLD(at, 2), ew = FLD(at, 3);
    float sx = (s.sum.x - h2 * ez) / s.diag, sy = (s.sum.y - h2 * ew) / s.diag;
    FLD(at, 0) = wmix(ex, sx, omega);
    FLD(at, 1) = wmix(ey, sy, omega);
}
static void k_smooth(OWorld *w, uint32_t l, int color) {
    PAR_FOR
    for (int32_t y = 0; y < (int32_t)w->lh[l]; y++)
        for (int32_t x = 0; x < (int32_t)w->lw[l]; x++) smoothAt(w, l, I2(x, y), color, 1.0f);
}
static void k_restrict(OWorld *w, uint32_t l) {              /* :2342-2377 */
    if (activeCount(w) == 0u || converged(w)) return;
    PAR_FOR
    for (int32_t y = 0; y < (int32_t)w->lh[l]; y++)
        for (int32_t x = 0; x < (int32_t)w->lw[l]; x++) {
            i2 c = I2(x, y);
            uint64_t at = levelSlot(w, l, c);
            int32_t span = 1 << l;
            if (span <= BLOCK) {
                i2 b = I2(c.x * span / BLOCK, c.y * span / BLOCK);
                if (inBlocks(w, b) && (w->blocks[STATE_SLOT(blockIndex(w, b))] & ACTIVE) == 0u) {
                    for (int j = 0; j < 4; j++) { FLD(at, j) = 0.0f; FLD(at + 1u, j) = 0.0f; }
                    continue;
                }
            }
            float fluid = 0.0f, sum = 0.0f;
            int air = 0;
            v2 r = V2(0.0f, 0.0f);
            for (int j = 0; j < 4; j++) {
                i2 k = I2(2 * c.x + (j & 1), 2 * c.y + (j >> 1));
                float rho = rhoAt(w, l - 1u, k);
                if (rho == 0.0f) air = MUTANT(3) ? 0 : 1;
                else if (rho > 0.0f) {
                    fluid += 1.0f;
                    sum += rho;
                    v2 q = residualAt(w, l - 1u, k);
                    r.x += q.x;
                    r.y += q.y;
                }
            }
            float rho = air ? 0.0f : (fluid > 0.0f ? sum / fluid : -1.0f);
            FLD(at, 0) = 0.0f;
            FLD(at, 1) = 0.0f;
            const float quarter = MUTANT(4) ? 1.0f / fluid : 0.25f;
            FLD(at, 2) = rho > 0.0f ? r.x * quarter : 0.0f;
            FLD(at, 3) = rho > 0.0f ? r.y * quarter : 0.0f;
            FLD(at + 1u, 0) = rho;
            FLD(at + 1u, 1) = 0.0f;
            FLD(at + 1u, 2) = 0.0f;
            FLD(at + 1u, 3) = 0.0f;
        }
}
static void k_coarsest(OWorld *w, uint32_t l) {              /* solve, :2407-2456 */
    if (converged(w)) return;
    uint32_t sx = w->lw[l], sy = w->lh[l];
    uint32_t n = sx * sy < MG_COARSEST ? sx * sy : MG_COARSEST;
    static float pot[MG_COARSEST][2], rhs[MG_COARSEST][2], rho[MG_COARSEST];
    for (uint32_t j = 0; j < n; j++) {
        uint64_t at = levelSlot(w, l, I2((int32_t)(j % sx), (int32_t)(j / sx)));
        pot[j][0] = FLD(at, 0); pot[j][1] = FLD(at, 1);
        rhs[j][0] = FLD(at, 2); rhs[j][1] = FLD(at, 3);
        rho[j] = FLD(at + 1u, 0);
    }
    float h2 = (float)(1u << (2u * l));
    uint32_t sweeps = w->ctl->coarse_sweeps;
    for (uint32_t s = 0; s < 2u * sweeps; s++) {
        for (uint32_t j = 0; j < n; j++) {
            i2 c = I2((int32_t)(j % sx), (int32_t)(j / sx));
            float r0 = rho[j];
            if ((uint32_t)((c.x + c.y) & 1) != (s & 1u) || r0 <= 0.0f) continue;
            float sumx = 0.0f, sumy = 0.0f, diag = 0.0f;
            for (int d = 0; d < 4; d++) {
                i2 o = beside(c, d);
                if (o.x < 0 || o.y < 0 || o.x >= (int32_t)sx || o.y >= (int32_t)sy) continue;
                uint32_t oj = (uint32_t)o.y * sx + (uint32_t)o.x;
                float r = rho[oj];
                if (r < 0.0f) continue;
                float beta = faceBeta(r0, r);
                sumx += beta * pot[oj][0];
                sumy += beta * pot[oj][1];
                diag += beta;
            }
            if (diag > 0.0f) {
                const float omega = MUTANT(6) ? 1.0f : OMEGA;
                pot[j][0] = wmix(pot[j][0], (sumx - h2 * rhs[j][0]) / diag, omega);
                pot[j][1] = wmix(pot[j][1], (sumy - h2 * rhs[j][1]) / diag, omega);
            }
        }
    }
    for (uint32_t j = 0; j < n; j++) {
        uint64_t at = levelSlot(w, l, I2((int32_t)(j % sx), (int32_t)(j / sx)));
        FLD(at, 0) = pot[j][0]; FLD(at, 1) = pot[j][1];
        FLD(at, 2) = rhs[j][0]; FLD(at, 3) = rhs[j][1];
    }
}
static v2 prolonged(const OWorld *w, uint32_t l, i2 c) {   /* :2478-2497 */
    i2 coarse = I2(c.x / 2, c.y / 2);
    i2 s = I2((c.x & 1) == 1 ? 1 : -1, (c.y & 1) == 1 ? 1 : -1);
    const i2 offsets[4] = {{0, 0}, {s.x, 0}, {0, s.y}, {s.x, s.y}};
    const float weights[4] = {9.0f, MUTANT(2) ? 9.0f : 3.0f, MUTANT(2) ? 9.0f : 3.0f, MUTANT(2) ? 9.0f : 1.0f};
    v2 sum = V2(0.0f, 0.0f);
    float weight = 0.0f;
    for (int k = 0; k < 4; k++) {
        i2 at = I2(coarse.x + offsets[k].x, coarse.y + offsets[k].y);
        float r = rhoAt(w, l + 1u, at);
        if (r < 0.0f) continue;
        weight += weights[k];
        if (r > 0.0f) {
            v2 p = potAt(w, l + 1u, at);
            sum.x += weights[k] * p.x;
            sum.y += weights[k] * p.y;
        }
    }
    float d = fmaxf(weight, 1e-6f);
    return V2(sum.x / d, sum.y / d);
}
static void k_prolong(OWorld *w, uint32_t l) {               /* :2459-2470 */
    if (activeCount(w) == 0u || converged(w)) return;
    PAR_FOR
    for (int32_t y = 0; y < (int32_t)w->lh[l]; y++)
        for (int32_t x = 0; x < (int32_t)w->lw[l]; x++) {
            uint64_t at = levelSlot(w, l, I2(x, y));
            if (FLD(at + 1u, 0) > 0.0f) {
                v2 e = prolonged(w, l, I2(x, y));
                FLD(at, 0) = FLD(at, 0) + e.x;
                FLD(at, 1) = FLD(at, 1) + e.y;
            }
        }
}
static void k_prolongFine(OWorld *w) {                       /* :2500-2512 */
    if (converged(w)) return;
    FOR_ACTIVE_CELLS(w, c) {
        if (!inWorld(w, c)) continue;
        uint32_t i = 2u * cellIndex(w, c);
        if (FLD(i, 2) > 0.0f) {
            v2 e = prolonged(w, 0u, c);
            FLD(i, 0) += e.x;
            FLD(i + 1u, 0) += e.y;
        }
    }
}

/* relaxRims + relaxTile (:2096-2192) */
static void k_relaxRims(OWorld *w) {
    if (converged(w)) return;
    FOR_ACTIVE_CELLS(w, c) {
        if (!inWorld(w, c)) continue;
        uint32_t i = cellIndex(w, c);
        if (FLD(2u * i, 2) > 0.0f) {
            CNT(i, G_RIM) = ubits(FLD(2u * i, 0));
            CNT(i, G_RIM + 1u) = ubits(FLD(2u * i + 1u, 0));
        }
    }
}
static i2 tileBlock(const OWorld *w, uint32_t wid, int shift) {
    if (wid < activeCount(w)) return activeBlock(w, wid);
    int32_t wide = (int32_t)w->bw, high = (int32_t)(w->B / w->bw);
    int32_t j = (int32_t)wid - (int32_t)w->B;
    if (shift == 0 || j < 0 || j > high + wide) return I2(-1, -1);
    return j <= high ? I2(wide, j) : I2(j - high - 1, high);
}
static void k_relax(OWorld *w, int shift) {
    if (converged(w)) return;
    uint32_t groups = shift ? w->B + w->bw + w->bh + 1u : w->B;
    PAR_FOR
    for (uint32_t wid = 0; wid < groups; wid++) {
        float pot[RIM * RIM][2], rho[RIM * RIM], share[RIM * RIM];
        i2 b = tileBlock(w, wid, shift);
        i2 origin = I2(b.x * BLOCK - shift, b.y * BLOCK - shift);
        if (b.x < 0 || origin.x >= w->W || origin.y >= w->H) continue;
        if ((statesOf(w, idiv8(origin), idiv8(I2(origin.x + BLOCK - 1, origin.y + BLOCK - 1))) & HAS_FLUID) == 0u) continue;
        for (int k = 0; k < RIM * RIM; k++) {
            i2 c = I2(origin.x - 1 + k % RIM, origin.y - 1 + k / RIM);
            float px = 0.0f, py = 0.0f, r = -1.0f;
            if (inWorld(w, c)) {
                uint32_t j = 2u * cellIndex(w, c);
                px = FLD(j, 0);
                py = FLD(j + 1u, 0);
                r = FLD(j, 2);
                if (r > 0.0f) {
                    px = fbits(CNT(cellIndex(w, c), G_RIM));
                    py = fbits(CNT(cellIndex(w, c), G_RIM + 1u));
                }
            }
            pot[k][0] = px;
            pot[k][1] = py;
            rho[k] = r;
            share[k] = densityAt(w, c);
        }
        float rhsx[64], rhsy[64], beta[64][4], diags[64][2], diag0[64];
        int solving[64], color[64];
        for (int li = 0; li < 64; li++) {
            int lx = li % BLOCK, ly = li / BLOCK;
            i2 me = I2(origin.x + lx, origin.y + ly);
            int t = (ly + 1) * RIM + lx + 1;
            float r0 = rho[t];
            rhsx[li] = rhsy[li] = 0.0f;
            beta[li][0] = beta[li][1] = beta[li][2] = beta[li][3] = 0.0f;
            if (r0 > 0.0f) {
                uint32_t i = 2u * cellIndex(w, me);
                rhsx[li] = FLD(i, 1);
                rhsy[li] = FLD(i + 1u, 1);
                float d = share[t];
                beta[li][0] = MUTANT(19) ? faceBeta(r0, rho[t - 1]) : fineBeta(r0, rho[t - 1], d, share[t - 1]);
                beta[li][1] = fineBeta(r0, rho[t + 1], d, share[t + 1]);
                beta[li][2] = fineBeta(r0, rho[t - RIM], d, share[t - RIM]);
                beta[li][3] = fineBeta(r0, rho[t + RIM], d, share[t + RIM]);
            }
            float diag = beta[li][0] + beta[li][1] + beta[li][2] + beta[li][3];
            int pinned = r0 > 0.0f && pinnedAt(w, cellIndex(w, me));
            diag0[li] = diag;
            diags[li][0] = diag + (pinned ? PIN : 0.0f);
            diags[li][1] = fmaxf(diag, 1e-30f);
            solving[li] = diags[li][0] > 0.0f;
            color[li] = (me.x + me.y) & 1;
        }
        for (int s = 0; s < 2 * (int)w->ctl->sweep_pairs; s++) {
            for (int li = 0; li < 64; li++) {
                if (!solving[li] || color[li] != (s & 1)) continue;
                int t = (li / BLOCK + 1) * RIM + li % BLOCK + 1;
                float sx = beta[li][0] * pot[t - 1][0] + beta[li][1] * pot[t + 1][0] + beta[li][2] * pot[t - RIM][0] + beta[li][3] * pot[t + RIM][0];
                float sy = beta[li][0] * pot[t - 1][1] + beta[li][1] * pot[t + 1][1] + beta[li][2] * pot[t - RIM][1] + beta[li][3] * pot[t + RIM][1];
                float nx = wmix(pot[t][0], (sx - rhsx[li]) / diags[li][0], SMOOTH_OMEGA);
                float ny = wmix(pot[t][1], (sy - rhsy[li]) / diags[li][1], SMOOTH_OMEGA);
                pot[t][0] = nx;
                pot[t][1] = diag0[li] > 0.0f ? ny : pot[t][1];
            }
        }
        for (int li = 0; li < 64; li++) {
            if (!solving[li]) continue;
            i2 me = I2(origin.x + li % BLOCK, origin.y + li / BLOCK);
            int t = (li / BLOCK + 1) * RIM + li % BLOCK + 1;
            uint32_t i = 2u * cellIndex(w, me);
            FLD(i, 0) = pot[t][0];
            FLD(i + 1u, 0) = pot[t][1];
        }
    }
}

/* --------------------------------------------------- conjugate gradients */
static float applyA(OWorld *w, i2 c, float rho) {            /* :2595-2613 */
    float sum = 0.0f, diag = 0.0f, share = densityAt(w, c);
    for (int d = 0; d < 4; d++) {
        i2 n = beside(c, d);
        float r = rhoAt(w, 0u, n);
        if (r < 0.0f) continue;
        float beta = fineBeta(rho, r, share, densityAt(w, n));
        if (r > 0.0f) sum += beta * CG(cellIndex(w, n), 2);
        diag += beta;
    }
    diag += pinnedAt(w, cellIndex(w, c)) ? PIN : 0.0f;
    return diag * CG(cellIndex(w, c), 2) - sum;
}

/* A block's partial sums, in the reference's f32 tree (:2628-2641) or in
 * FP64 with the same tree. */
static void sumBlockF(float v[64][2], float out[2]) {
    for (int h = 32; h > 0; h >>= 1)
        for (int li = 0; li < h; li++) { v[li][0] += v[li + h][0]; v[li][1] += v[li + h][1]; }
    out[0] = v[0][0];
    out[1] = v[0][1];
}
static void sumBlockD(double v[64][3], double out[3]) {
    for (int h = 32; h > 0; h >>= 1)
        for (int li = 0; li < h; li++) { v[li][0] += v[li + h][0]; v[li][1] += v[li + h][1]; v[li][2] += v[li + h][2]; }
    out[0] = v[0][0];
    out[1] = v[0][1];
    out[2] = v[0][2];
}
/* The active blocks' partials summed in block-index order by 256 threads
 * (:2645-2661). */
static void sumBlocksF(OWorld *w, float out[2]) {
    float s[256][2];
    for (int li = 0; li < 256; li++) {
        s[li][0] = s[li][1] = 0.0f;
        for (uint32_t k = (uint32_t)li; k < w->B; k += 256u)
            if (w->blocks[STATE_SLOT(k)] & ACTIVE) { s[li][0] += CG(w->C + k, 0); s[li][1] += CG(w->C + k, 1); }
    }
    for (int h = 128; h > 0; h >>= 1)
        for (int li = 0; li < h; li++) { s[li][0] += s[li + h][0]; s[li][1] += s[li + h][1]; }
    out[0] = s[0][0];
    out[1] = s[0][1];
}
/* FP64: the partials are kept by the block's place in the active list (in
 * block-index order), and summed by 1024 lanes, lane i taking places i,
 * i + 1024, ..., then halving. */
static void sumBlocksD(OWorld *w, int which, double *out) {
    double s[1024];
    double *bp = bpart(w);
    uint32_t na = activeCount(w);
    for (int li = 0; li < 1024; li++) {
        s[li] = 0.0;
        for (uint32_t j = (uint32_t)li; j < na; j += 1024u) s[li] += bp[4ull * j + (uint32_t)which];
    }
    for (int h = 512; h > 0; h >>= 1)
        for (int li = 0; li < h; li++) s[li] += s[li + h];
    *out = s[0];
}

static void k_cgStart(OWorld *w) {
    FOR_ACTIVE_CELLS(w, c) {
        if (!inWorld(w, c)) continue;
        uint32_t i = cellIndex(w, c);
        CG(i, 0) = gridX(w, c);
        CG(i, 1) = 0.0f;
        CG(i, 2) = 0.0f;
        CG(i, 3) = 0.0f;
        CNT(i, G_REGION) = i;
    }
}
static void k_cgJoin(OWorld *w) {
    FOR_ACTIVE_CELLS_SERIAL(w, c) {
        if (!inWorld(w, c) || gridZ(w, c) <= 0.0f) continue;
        for (int d = 0; d < 2; d++) {
            i2 n = d == 0 ? I2(c.x + 1, c.y) : I2(c.x, c.y + 1);
            if (rhoAt(w, 0u, n) > 0.0f) ufJoin(w, cellIndex(w, c), cellIndex(w, n));
        }
    }
}
static void k_cgAnchor(OWorld *w) {
    FOR_ACTIVE_CELLS_SERIAL(w, c) {
        if (!inWorld(w, c) || gridZ(w, c) <= 0.0f) continue;
        for (int d = 0; d < 4; d++) {
            i2 n = beside(c, d);
            if (inWorld(w, n) && rhoAt(w, 0u, n) == 0.0f) {
                CNT(ufFind(w, cellIndex(w, c)), G_REGION) |= REGION_OPEN;
                break;
            }
        }
    }
}
static void k_cgResidual(OWorld *w) {
    double *bp = bpart(w);
    PAR_FOR
    for (uint32_t wid = 0; wid < activeCount(w); wid++) {
        i2 b = activeBlock(w, wid);
        double v[64][3];
        for (int li = 0; li < 64; li++) {
            i2 c = I2(b.x * BLOCK + li % BLOCK, b.y * BLOCK + li / BLOCK);
            v[li][0] = v[li][1] = v[li][2] = 0.0;
            if (!inWorld(w, c) || gridZ(w, c) <= 0.0f) continue;
            uint32_t i = cellIndex(w, c);
            CG(i, 1) = -residualAt(w, 0u, c).x - couplingAt(w, c, PUSH_X);   /* (:2714-2716) */
            double bi = -(double)FLD(2u * i, 1);
            v[li][2] = bi * bi;
            if (w->capture && w->cap_b) w->cap_b[i] = -FLD(2u * i, 1);
            w->bvec[i] = -FLD(2u * i, 1);
        }
        double out[3];
        sumBlockD(v, out);
        bp[4ull * wid + 2] = out[2];
    }
    double bb;
    sumBlocksD(w, 2, &bb);
    head(w)->bb = bb;
}
static void k_cgPrepare(OWorld *w) {
    FOR_ACTIVE_CELLS(w, c) {
        if (!inWorld(w, c)) continue;
        uint32_t i = cellIndex(w, c);
        FLD(2u * i, 0) = 0.0f;
        if (FLD(2u * i, 2) > 0.0f) FLD(2u * i, 1) = -CG(i, 1);
    }
}
/* cgRho + cgBeta (:2736-2764) */
static void k_cgRhoBeta(OWorld *w) {
    if (converged(w)) return;
    float *sums = &CG(w->C + w->B, 0);
    if (w->dot_mode == 0) {
        PAR_FOR
        for (uint32_t wid = 0; wid < activeCount(w); wid++) {
            i2 b = activeBlock(w, wid);
            float v[64][2];
            for (int li = 0; li < 64; li++) {
                i2 c = I2(b.x * BLOCK + li % BLOCK, b.y * BLOCK + li / BLOCK);
                v[li][0] = v[li][1] = 0.0f;
                if (inWorld(w, c) && gridZ(w, c) > 0.0f) {
                    uint32_t i = cellIndex(w, c);
                    float z = FLD(2u * i, 0);
                    v[li][0] = CG(i, 1) * z;
                    v[li][1] = CG(i, 3) * z;
                }
            }
            float out[2];
            sumBlockF(v, out);
            CG(w->C + blockIndex(w, b), 0) = out[0];
            CG(w->C + blockIndex(w, b), 1) = out[1];
            CG(w->C + blockIndex(w, b), 2) = 0.0f;
            CG(w->C + blockIndex(w, b), 3) = 0.0f;
        }
        float s[2];
        sumBlocksF(w, s);
        float wasw = sums[3];
        float beta = (w->cycle == 0u || wasw <= 0.0f) ? 0.0f : -s[1] / wasw;
        sums[0] = s[0];
        sums[2] = beta;
        head(w)->rz = s[0];
        head(w)->beta = beta;
    } else {
        double *bp = bpart(w);
        PAR_FOR
        for (uint32_t wid = 0; wid < activeCount(w); wid++) {
            i2 b = activeBlock(w, wid);
            double v[64][3];
            for (int li = 0; li < 64; li++) {
                i2 c = I2(b.x * BLOCK + li % BLOCK, b.y * BLOCK + li / BLOCK);
                v[li][0] = v[li][1] = v[li][2] = 0.0;
                if (inWorld(w, c) && gridZ(w, c) > 0.0f) {
                    uint32_t i = cellIndex(w, c);
                    double z = (double)FLD(2u * i, 0);
                    v[li][0] = (double)CG(i, 1) * z;
                    v[li][1] = (double)CG(i, 3) * z;
                }
            }
            double out[3];
            sumBlockD(v, out);
            bp[4ull * wid + 0] = out[0];
            bp[4ull * wid + 1] = out[1];
        }
        double s0, s1;
        sumBlocksD(w, 0, &s0);
        sumBlocksD(w, 1, &s1);
        SolveHead *h = head(w);
        double beta = (w->cycle == 0u || h->pq <= 0.0) ? 0.0 : -s1 / h->pq;
        h->rz = s0;
        h->beta = beta;
        sums[0] = (float)s0;
        sums[2] = (float)beta;
    }
}
static void k_cgDirection(OWorld *w) {                       /* :2767-2776 */
    if (converged(w)) return;
    float betaf = CG(w->C + w->B, 2);
    double betad = head(w)->beta;
    FOR_ACTIVE_CELLS(w, c) {
        if (!inWorld(w, c)) continue;
        uint32_t i = cellIndex(w, c);
        float p;
        if (FLD(2u * i, 2) > 0.0f)
            p = w->dot_mode == 0 ? FLD(2u * i, 0) + betaf * CG(i, 2) : (float)((double)FLD(2u * i, 0) + betad * (double)CG(i, 2));
        else
            p = 0.0f;
        CG(i, 2) = p;
    }
}
/* cgApply + cgAlpha (:2779-2814) */
static void k_cgApplyAlpha(OWorld *w) {
    if (converged(w)) return;
    float *sums = &CG(w->C + w->B, 0);
    double *bp = bpart(w);
    PAR_FOR
    for (uint32_t wid = 0; wid < activeCount(w); wid++) {
        i2 b = activeBlock(w, wid);
        float vf[64][2];
        double vd[64][3];
        for (int li = 0; li < 64; li++) {
            i2 c = I2(b.x * BLOCK + li % BLOCK, b.y * BLOCK + li / BLOCK);
            vf[li][0] = vf[li][1] = 0.0f;
            vd[li][0] = vd[li][1] = vd[li][2] = 0.0;
            if (!inWorld(w, c)) continue;
            float rho = gridZ(w, c);
            uint32_t i = cellIndex(w, c);
            float q = rho > 0.0f ? applyA(w, c, rho) + couplingAt(w, c, PUSH_P0 + w->cycle % 2u) : 0.0f;   /* (:2790) */
            CG(i, 3) = q;
            vf[li][0] = CG(i, 2) * q;
            vd[li][0] = (double)CG(i, 2) * (double)q;
        }
        if (w->dot_mode == 0) {
            float out[2];
            sumBlockF(vf, out);
            CG(w->C + blockIndex(w, b), 0) = out[0];
            CG(w->C + blockIndex(w, b), 1) = out[1];
        } else {
            double out[3];
            sumBlockD(vd, out);
            bp[4ull * wid + 0] = out[0];
        }
    }
    if (w->dot_mode == 0) {
        float s[2];
        sumBlocksF(w, s);
        float pq = s[0];
        sums[1] = pq <= 0.0f ? 0.0f : sums[0] / pq;
        sums[3] = pq;
        head(w)->alpha = sums[1];
        head(w)->pq = pq;
    } else {
        double pq;
        sumBlocksD(w, 0, &pq);
        SolveHead *h = head(w);
        h->alpha = pq <= 0.0 ? 0.0 : h->rz / pq;
        h->pq = pq;
        sums[1] = (float)h->alpha;
        sums[3] = (float)pq;
    }
    /* the bodies' sums of this iteration's direction, done with, emptied
     * for the iteration after next (:2808-2813) */
    if (bodiesOn(w))
        for (uint32_t k = 0; k < MAX_BODIES * 6u; k++) w->pushes[pushSlot(k / 6u, PUSH_P0 + w->cycle % 2u, 0u) + k % 6u] = 0u;
}
/* cgUpdate (:2852-2880) and the convergence test the next iteration reads */
static void k_cgUpdate(OWorld *w) {
    w->blocks[COUNT_SLOT(COUNT_RESIDUAL + (w->cycle + 1u) % 3u)] = 0u;
    if (!converged(w)) w->blocks[COUNT_SLOT(COUNT_CYCLES)] += 1u;
    if (converged(w)) return;
    float alphaf = CG(w->C + w->B, 1);
    double alphad = head(w)->alpha;
    double *bp = bpart(w);
    uint32_t most = 0;
    PAR_FOR_MOST
    for (uint32_t wid = 0; wid < activeCount(w); wid++) {
        i2 b = activeBlock(w, wid);
        double v[64][3];
        for (int li = 0; li < 64; li++) {
            i2 c = I2(b.x * BLOCK + li % BLOCK, b.y * BLOCK + li / BLOCK);
            v[li][0] = v[li][1] = v[li][2] = 0.0;
            if (!inWorld(w, c) || gridZ(w, c) <= 0.0f) continue;
            uint32_t i = cellIndex(w, c);
            float x = CG(i, 0), r = CG(i, 1), p = CG(i, 2), q = CG(i, 3), rn, xn;
            if (w->dot_mode == 0) {
                rn = r - alphaf * q;
                xn = x + alphaf * p;
            } else {
                rn = (float)((double)r - alphad * (double)q);
                xn = (float)((double)x + alphad * (double)p);
            }
            CG(i, 0) = xn;
            CG(i, 1) = rn;
            FLD(2u * i, 0) = 0.0f;
            FLD(2u * i, 1) = -rn;
            uint32_t bits = ubits(fabsf(rn));
            if (bits > most) most = bits;
            v[li][2] = (double)rn * (double)rn;
        }
        double out[3];
        sumBlockD(v, out);
        bp[4ull * wid + 2] = out[2];
    }
    /* residual replacement (WigStepCtl.residual_every, the C ABI's): after
     * every k-th iteration r = b - A x from the updated x, the products
     * summed in FP64 in beside()'s order, in place of the recursive r */
    const uint32_t every = w->ctl->residual_every;
    const int replace = every != 0u && (w->cycle + 1u) % every == 0u;
    if (replace) {
        most = 0;
        PAR_FOR_MOST
        for (uint32_t wid = 0; wid < activeCount(w); wid++) {
            i2 b = activeBlock(w, wid);
            double v[64][3];
            for (int li = 0; li < 64; li++) {
                i2 c = I2(b.x * BLOCK + li % BLOCK, b.y * BLOCK + li / BLOCK);
                v[li][0] = v[li][1] = v[li][2] = 0.0;
                if (!inWorld(w, c) || gridZ(w, c) <= 0.0f) continue;
                uint32_t i = cellIndex(w, c);
                float rho = gridZ(w, c), share = densityAt(w, c);
                double sum = 0.0, diag = 0.0, mag = 0.0;
                for (int d = 0; d < 4; d++) {
                    i2 n = beside(c, d);
                    float r = rhoAt(w, 0u, n);
                    if (r < 0.0f) continue;
                    float beta = fineBeta(rho, r, share, densityAt(w, n));
                    if (beta == 0.0f) continue;
                    if (r > 0.0f) {
                        double xn = (double)CG(cellIndex(w, n), 0);
                        sum += (double)beta * xn;
                        mag += (double)beta * fabs(xn);
                    }
                    diag += (double)beta;
                }
                if (pinnedAt(w, i)) diag += (double)PIN;
                float rn = (float)((double)w->bvec[i] - (diag * (double)CG(i, 0) - sum));
                CG(i, 1) = rn;
                FLD(2u * i, 1) = -rn;
                uint32_t bits = ubits(fabsf(rn));
                if (bits > most) most = bits;
                v[li][2] = (double)rn * (double)rn;
                double a = diag * fabs((double)CG(i, 0)) + mag;    /* (|A| |x|) at i */
                v[li][1] = a * a;
            }
            double out[3];
            sumBlockD(v, out);
            bp[4ull * wid + 1] = out[1];
            bp[4ull * wid + 2] = out[2];
        }
    }
    uint32_t *slot = &w->blocks[COUNT_SLOT(COUNT_RESIDUAL + w->cycle % 3u)];
    if (most > *slot) *slot = most;
    /* cgCheck: whether the next iteration is skipped */
    SolveHead *h = head(w);
    h->cycles += 1u;
    if (w->ctl->tol_mode == WIG_TOL_REL_L2) {
        double rr;
        sumBlocksD(w, 2, &rr);
        h->rr = rr;
        double t = (double)w->ctl->tol;
        w->converged_flag = rr <= t * t * h->bb;
    } else {
        double rr;
        sumBlocksD(w, 2, &rr);
        h->rr = rr;
        w->converged_flag = fbits(*slot) < w->ctl->tol;
    }
    /* with residual replacement, only a replaced (true) residual stops the
     * solve: under the tolerance, or within twice the rounding floor of an
     * f32 iterate, ||b - A x|| <= 2^-25 ||(|A| |x|)|| (the floor measures
     * about 2^-24 / 4 of it) */
    if (every != 0u) {
        double aa = 0.0;
        if (replace) sumBlocksD(w, 1, &aa);
        w->converged_flag = replace && (w->converged_flag || h->rr <= 8.8817841970012523e-16 * aa);
    }
    h->converged = (uint32_t)w->converged_flag;
}
static void k_cgFinish(OWorld *w) {
    FOR_ACTIVE_CELLS(w, c) {
        if (!inWorld(w, c)) continue;
        uint32_t i = cellIndex(w, c);
        FLD(2u * i, 0) = CG(i, 0);
        CNT(i, G_REGION) = 0u;
    }
}

/* --------------------------------------------------------------- project */
static void projectFace(OWorld *w, i2 at, int isU) {        /* :2897-2923 */
    uint64_t f = faceSlot(w, faceIndex(w, at, isU));
    i2 a = isU ? I2(at.x - 1, at.y) : I2(at.x, at.y - 1);
    float ga[4], gb[4];
    gridAt4(w, a, ga);
    gridAt4(w, at, gb);
    float beta = fineBeta(ga[2], gb[2], densityAt(w, a), densityAt(w, at));
    uint32_t ra = rigidAt(w, a), rb = rigidAt(w, at);
    if (beta > 0.0f) {
        FLD(f, 0) = MUTANT(13) ? FLD(f, 0) + beta * (gb[0] - ga[0]) : FLD(f, 0) - beta * (gb[0] - ga[0]);
        FLD(f, 2) = PROJECTED;
        /* no drift into or out of a body's cells (:2907-2909) */
        FLD(f, 3) = ra != NONE || rb != NONE ? 0.0f : wclamp(-beta * (driftAt(w, at) - driftAt(w, a)), -MAX_DRIFT, MAX_DRIFT);
    }
    /* a face between the liquid and a free body moves as the body does with
     * the pressure's push the solve found (:2911-2921) */
    if ((ra == NONE) != (rb == NONE) && (ra != NONE ? gb[2] : ga[2]) > 0.0f) {
        uint32_t b = ra != NONE ? ra : rb;
        v3 ch = pushChange(w, b, PUSH_F);
        v2 m = faceMiddle(at, isU), pos = bodyV(w, b, BW_POS);
        v2 r = V2(m.x - pos.x, m.y - pos.y);
        FLD(f, 0) += isU ? ch.x - ch.z * r.y : ch.y + ch.z * r.x;
        FLD(f, 2) = PROJECTED;
    }
}
static void k_project(OWorld *w) {
    uint32_t *most = &w->blocks[COUNT_SLOT(COUNT_MOST_CYCLES)];
    uint32_t cyc = w->blocks[COUNT_SLOT(COUNT_CYCLES)];
    if (cyc > *most) *most = cyc;
    FOR_ACTIVE_CELLS(w, c) {
        if (!inWorld(w, c)) continue;
        Face3 fs[4];
        facesOf(w, c, fs);
        for (int j = 0; j < 4; j++)
            if (fs[j].isU >= 0) projectFace(w, I2(fs[j].x, fs[j].y), fs[j].isU == 1);
    }
}
static void extrapolateFace(OWorld *w, i2 at, int isU, int layer) {   /* :3226-3252 */
    uint64_t f = faceSlot(w, faceIndex(w, at, isU));
    i2 a = isU ? I2(at.x - 1, at.y) : I2(at.x, at.y - 1);
    if (FLD(f, 2) != 1.0f || gridZ(w, a) != 0.0f || gridZ(w, at) != 0.0f) return;
    float reached = layer == 2 && !MUTANT(12) ? 1.5f : PROJECTED;
    float sum = 0.0f, n = 0.0f;
    for (int d = 0; d < 4; d++) {
        i2 q = beside(at, d);
        if (!faceInWorld(w, q, isU)) continue;
        uint64_t g = faceSlot(w, faceIndex(w, q, isU));
        if (FLD(g, 2) >= reached) {
            sum += FLD(g, 0);
            n += 1.0f;
        }
    }
    if (n > 0.0f) {                        /* written in place: a face written here holds z < reached */
        FLD(f, 0) = sum / n;
        FLD(f, 2) = layer == 2 ? 1.25f : 1.5f;
    }
}
static void k_extrapolate(OWorld *w, int layer) {            /* :3254-3275 */
    FOR_ACTIVE_CELLS(w, c) {
        if (!inWorld(w, c)) continue;
        Face3 fs[4];
        facesOf(w, c, fs);
        for (int j = 0; j < 4; j++)
            if (fs[j].isU >= 0) extrapolateFace(w, I2(fs[j].x, fs[j].y), fs[j].isU == 1, layer);
    }
}

/* ---------------------------------------------------------------- advect */
typedef struct { float now, before, weight, drift; v2 gradient; } Sample;
static Sample sampleAxis(OWorld *w, v2 s, int isU) {         /* :3480-3512 */
    float fx0 = floorf(s.x), fy0 = floorf(s.y);
    i2 b = I2(wi32(fx0), wi32(fy0));
    v2 f = V2(s.x - fx0, s.y - fy0);
    float ax = 0.0f, ay = 0.0f;
    Sample r = {0.0f, 0.0f, 0.0f, 0.0f, {0.0f, 0.0f}};
    int held = 0;
    v2 g = V2(0.0f, 0.0f);
    for (int j = 0; j < 4; j++) {
        int ox = j & 1, oy = j >> 1;
        float wgt = (ox == 1 ? f.x : 1.0f - f.x) * (oy == 1 ? f.y : 1.0f - f.y);
        v2 dw = V2((ox == 1 ? 1.0f : -1.0f) * (oy == 1 ? f.y : 1.0f - f.y), (ox == 1 ? f.x : 1.0f - f.x) * (oy == 1 ? 1.0f : -1.0f));
        i2 at = I2(b.x + ox, b.y + oy);
        if (!faceInWorld(w, at, isU)) continue;
        uint64_t fs = faceSlot(w, faceIndex(w, at, isU));
        float fx = FLD(fs, 0), fy = FLD(fs, 1), fz = FLD(fs, 2), fw = FLD(fs, 3);
        r.drift += wgt * fw;
        if (fz > 0.0f) {
            ax += wgt * fx;
            ay += wgt * fy;
            r.weight += wgt;
            g.x += dw.x * fx;
            g.y += dw.y * fx;
            held += fz == PROJECTED ? 1 : 0;
        }
    }
    if (r.weight > 0.0f) {
        r.now = ax / r.weight;
        r.before = ay / r.weight;
    }
    r.gradient = held == 4 ? g : V2(0.0f, 0.0f);
    return r;
}
static inline v2 clampSpeed(v2 v) {                           /* :437-440 */
    const float most = MUTANT(7) ? 16.0f : LIQUID_MAX_SPEED;
    float s = wlength(v);
    return s > most ? V2(v.x * (most / s), v.y * (most / s)) : v;
}
static int escapesTo(OWorld *w, i2 q, uint32_t side) {       /* :3535-3541 */
    if (!inWorld(w, q) || shutAt(w, q)) return 0;
    uint32_t cup = cupOf(w, cellIndex(w, q));
    return cup == NONE || cup == side;
}
static v2 escape(OWorld *w, v2 pos) {                        /* :3552-3594 */
    if (!isObstaclePos(w, pos) || MUTANT(22)) return pos;
    i2 c = I2(wi32(floorf(pos.x)), wi32(floorf(pos.y)));
    uint32_t side = inWorld(w, c) ? cupBefore(w, cellIndex(w, c)) : NONE;
    v2 best = pos;
    float bestD = 1e9f;
    for (int n = 0; n < 8; n++) {
        int m = n + (n >= 4 ? 1 : 0);
        i2 at = I2(c.x + m % 3 - 1, c.y + m / 3 - 1);
        if (!escapesTo(w, at, side)) continue;
        v2 inside = V2(wclamp(pos.x, (float)at.x + 0.02f, (float)at.x + 0.98f), wclamp(pos.y, (float)at.y + 0.02f, (float)at.y + 0.98f));
        float d = wlength(V2(pos.x - inside.x, pos.y - inside.y));
        if (d < bestD) { bestD = d; best = inside; }
    }
    if (bestD < 1e9f) return best;
    static const i2 ways[4] = {{0, -1}, {-1, 0}, {1, 0}, {0, 1}};
    int32_t k = downTurns(w, c);
    i2 way = turnStep(ways[0], k);
    int nearest = ESCAPE_REACH + 1;
    for (int n = 0; n < 4; n++) {
        i2 dir = turnStep(ways[n], k);
        for (int r = 2; r < nearest; r++) {
            i2 q = I2(c.x + dir.x * r, c.y + dir.y * r);
            if (!isObstaclePos(w, V2((float)q.x + 0.5f, (float)q.y + 0.5f)) && escapesTo(w, q, side)) {
                nearest = r;
                way = dir;
                break;
            }
        }
    }
    return inWorld(w, I2(c.x + way.x, c.y + way.y)) ? V2(pos.x + (float)way.x, pos.y + (float)way.y) : pos;
}
static v2 slipOf(OWorld *w, const Particle *q) {             /* :3605-3614 */
    uint32_t i = cellIndex(w, cellOf(w, q->pos));
    uint32_t n = CNT(i, G_COUNT);
    if (n < 2u) return V2(0.0f, 0.0f);
    float mean = (float)CNT(i, G_MASS) / (FIX * (float)n);
    float lighter = (mean - liquidDensity(q->word, q->temp)) / fmaxf(mean, 1e-3f);
    v2 pull = pullAt(w, q->pos);
    return V2(-pull.x / WATER_WEIGHT * SLIP * lighter, -pull.y / WATER_WEIGHT * SLIP * lighter);
}
static void k_advect(OWorld *w) {                            /* :3616-3694 */
    uint32_t n = pslots(w, w->parts);
    uint32_t *pb = w->parts;
    PAR_FOR
    for (uint32_t k = 0; k < n; k++) {
        Particle q = loadP(w, pb, k);
        v2 orig = q.pos;
        q.pos = escape(w, q.pos);
        Sample u = sampleAxis(w, V2(q.pos.x - 0.0f, q.pos.y - 0.5f), 1);
        Sample v = sampleAxis(w, V2(q.pos.x - 0.5f, q.pos.y - 0.0f), 0);
        v2 flow = V2(u.weight > 0.0f ? u.now : q.vel.x, v.weight > 0.0f ? v.now : q.vel.y);
        v2 vel = clampSpeed(flow);
        const float keep = MUTANT(1) ? 1.0f : AFFINE_KEEP;
        float affine[4] = {keep * u.gradient.x, keep * u.gradient.y, keep * v.gradient.x, keep * v.gradient.y};
        v2 cs = clampSpeed(vel), slip = slipOf(w, &q);
        v2 travel = V2(cs.x + wclamp(u.drift + slip.x, -MAX_DRIFT, MAX_DRIFT), cs.y + wclamp(v.drift + slip.y, -MAX_DRIFT, MAX_DRIFT));
        int32_t steps = wi32(ceilf(fmaxf(fabsf(travel.x), fabsf(travel.y)) * 2.0f));
        steps = steps < 1 ? 1 : (steps > ADVECT_STEPS ? ADVECT_STEPS : steps);
        v2 d = V2(travel.x / (float)steps, travel.y / (float)steps);
        v2 pos = q.pos;
        for (int32_t s = 0; s < steps; s++) {
            v2 nx = V2(pos.x + d.x, pos.y + 0.0f);
            if (d.x != 0.0f && isObstaclePos(w, nx)) {
                if (vel.x * d.x > 0.0f && !MUTANT(8)) vel.x = 0.0f;
                d.x = 0.0f;
            } else {
                pos = nx;
            }
            v2 ny = V2(pos.x + 0.0f, pos.y + d.y);
            if (d.y != 0.0f && isObstaclePos(w, ny)) {
                if (vel.y * d.y > 0.0f) vel.y = 0.0f;
                d.y = 0.0f;
            } else {
                pos = ny;
            }
        }
        if (!isActiveBlock(w, idiv8(cellOf(w, pos)))) {
            pos = q.pos;
            if (!isActiveBlock(w, idiv8(cellOf(w, pos)))) pos = orig;
        }
        MOTION(pb, k, 0) = ubits(pos.x);
        MOTION(pb, k, 1) = ubits(pos.y);
        MOTION(pb, k, 2) = ubits(vel.x);
        MOTION(pb, k, 3) = ubits(vel.y);
        for (int j = 0; j < 4; j++) AFFINE(pb, k, j) = ubits(affine[j]);
        uint32_t li = 2u * cellIndex(w, cellOf(w, q.pos));
        float lx = FLD(li, 0), lz = FLD(li, 2);
        if (lz > 0.0f) {
            int falling = lx < 0.5f * lz * wlength(pullAt(w, q.pos));
            at_or(&CNT(cellIndex(w, cellOf(w, pos)), G_CAME), falling ? CAME_FALLING : CAME_PRESSING);
        }
    }
}

/* ------------------------------------------------------------------ sort */
static void k_sortCount(OWorld *w) {                         /* :3700-3708 */
    uint32_t n = pslots(w, w->parts);
    PAR_FOR
    for (uint32_t k = 0; k < n; k++) {
        v2 pos = V2(fbits(MOTION(w->parts, k, 0)), fbits(MOTION(w->parts, k, 1)));
        STATE(w->parts, k, 3) = at_take(&CNT(cellIndex(w, cellOf(w, pos)), G_SORT), 1u);
    }
}
static void k_scanBlocks(OWorld *w) {                        /* :3730-3753 */
    PAR_FOR
    for (uint32_t wid = 0; wid < activeCount(w); wid++) {
        i2 b = activeBlock(w, wid);
        uint32_t run = 0;
        for (int li = 0; li < 64; li++) {
            i2 c = I2(b.x * BLOCK + li % BLOCK, b.y * BLOCK + li / BLOCK);
            uint32_t n = 0;
            if (inWorld(w, c)) n = CNT(cellIndex(w, c), G_SORT);
            run += n;
            if (inWorld(w, c)) {
                uint32_t i = cellIndex(w, c);
                CNT(i, G_START) = MUTANT(9) ? run : run - n;
                uint32_t came = CNT(i, G_CAME);
                CNT(i, G_CAME) = 0u;
                if (came == CAME_FALLING) FLD(2u * i + 1u, 3) = (float)(wu32(FLD(2u * i + 1u, 3)) | fallingBit(w, 0u));
            }
        }
        w->blocks[TOTAL_SLOT(wid)] = run;
    }
}
static void k_scanTotals(OWorld *w) {                        /* :3757-3780 */
    uint32_t n = activeCount(w), run = 0;
    for (uint32_t j = 0; j < n; j++) {
        uint32_t t = w->blocks[TOTAL_SLOT(j)];
        w->blocks[TOTAL_SLOT(j)] = run;
        run += t;
    }
    w->sorted[0] = run;
    w->blocks[COUNT_SLOT(COUNT_PARTICLES)] = run;
}
static void k_scanFinish(OWorld *w) {                        /* :3783-3790 */
    PAR_FOR
    for (uint32_t wid = 0; wid < activeCount(w); wid++) {
        i2 b = activeBlock(w, wid);
        for (int li = 0; li < 64; li++) {
            i2 c = I2(b.x * BLOCK + li % BLOCK, b.y * BLOCK + li / BLOCK);
            if (inWorld(w, c)) CNT(cellIndex(w, c), G_START) += w->blocks[TOTAL_SLOT(wid)];
        }
    }
}
static void k_scatter(OWorld *w) {                           /* :3793-3808 */
    uint32_t n = pslots(w, w->parts);
    PAR_FOR
    for (uint32_t k = 0; k < n; k++) {
        v2 pos = V2(fbits(MOTION(w->parts, k, 0)), fbits(MOTION(w->parts, k, 1)));
        uint32_t i = cellIndex(w, cellOf(w, pos));
        uint32_t m = CNT(i, G_START) + STATE(w->parts, k, 3);
        if (m < w->cap) {
            for (int j = 0; j < 4; j++) {
                MOTION(w->sorted, m, j) = MOTION(w->parts, k, j);
                STATE(w->sorted, m, j) = STATE(w->parts, k, j);
                AFFINE(w->sorted, m, j) = AFFINE(w->parts, k, j);
            }
        }
    }
}

/* ------------------------------------------------------------------ draw */
static inline int vacated(OWorld *w, uint32_t j) {
    return kindOf(w, wType(w->cells[4ull * j])) == KIND_LIQUID && CNT(j, G_SORT) == 0u;
}
static uint32_t airTo(OWorld *w, i2 c, int gas) {            /* :3898-3924 */
    if (!gas && gridZ(w, c) > 0.0f) return NONE;
    static const i2 airOrder[8] = {{0, -1}, {-1, 0}, {1, 0}, {0, 1}, {-1, -1}, {1, -1}, {-1, 1}, {1, 1}};
    static const i2 gasOrder[8] = {{0, -1}, {-1, -1}, {1, -1}, {-1, 0}, {1, 0}, {-1, 1}, {1, 1}, {0, 1}};
    int32_t k = downTurns(w, c);
    for (int n = 0; n < 8; n++) {
        i2 o = turnStep(gas ? gasOrder[n] : airOrder[n], k);
        i2 at = I2(c.x + o.x, c.y + o.y);
        if (!inWorld(w, at)) continue;
        uint32_t j = cellIndex(w, at);
        if (vacated(w, j) || (gas && wType(w->cells[4ull * j]) == E_AIR && CNT(j, G_SORT) == 0u && !isOccupant(occupantAt(w, j)))) return j;
    }
    return NONE;
}
static void k_rasterClaims(OWorld *w) {                      /* :3928-3948 */
    FOR_ACTIVE_CELLS(w, c) {
        if (!inWorld(w, c)) continue;
        uint32_t i = cellIndex(w, c), to = NONE;
        uint32_t t = wType(w->cells[4ull * i]);
        if (CNT(i, G_SORT) > 0u && !isOccupant(occupantAt(w, i))) {
            if (t == E_AIR) to = airTo(w, c, 0);
            else if (kindOf(w, t) == KIND_GAS && gridZ(w, c) >= 0.0f) to = airTo(w, c, 1);
        }
        CNT(i, G_AIR_TO) = to;
        if (to != NONE) at_max(&CNT(to, G_TAKEN), ~i);
    }
}
static int moveOutOfTheWay(OWorld *w, uint32_t i, const uint32_t v[4]) {   /* :3954-3973 */
    uint32_t to = CNT(i, G_AIR_TO);
    if (to != NONE && CNT(to, G_TAKEN) == ~i) {
        /* pushGasOut at the air's temperature as stored (AIR has no phase
         * points: js/elements.js:424-426), or, without the air, let go */
        if (wType(w->cells[4ull * to]) == E_AIR) {
            if (w->ctl->air) pushGasOut(w, to, fbits(w->cells[4ull * to + 1]));
            else setGasPair(w, to, 0u, 0u);
        }
        for (int j = 0; j < 4; j++) w->cells[4ull * to + j] = v[j];
        uint64_t gi = gasWord(w, i);
        setGasPair(w, to, w->cells[gi], w->cells[gi + 1]);
        setGasPair(w, i, 0u, 0u);
        return 1;
    }
    if (wType(v[0]) == E_AIR) {
        if (w->ctl->air) pushGasOut(w, i, fbits(v[1]));
        else setGasPair(w, i, 0u, 0u);
        return 1;
    }
    return 0;
}
static void k_raster(OWorld *w) {                            /* :3983-4119 */
    FOR_ACTIVE_CELLS(w, c) {
        if (!inWorld(w, c)) continue;
        uint32_t i = cellIndex(w, c), n = CNT(i, G_SORT);
        if (n == 0u) continue;
        uint32_t start = CNT(i, G_START);
        uint32_t v[4];
        for (int j = 0; j < 4; j++) v[j] = w->cells[4ull * i + j];
        uint32_t t = wType(v[0]);
        /* a person's body (:3995-4004): not drawn, its particles left to find
         * their way out, and liquid drawn there before the person came is air
         * (at the liquid's wordTempOf, water's effective temperature) */
        if (isOccupant(occupantAt(w, i))) {
            if (kindOf(w, t) == KIND_LIQUID && !MUTANT(29)) {
                w->cells[4ull * i] = E_AIR;
                w->cells[4ull * i + 1] = ubits(effTemp(fbits(v[1])));
                w->cells[4ull * i + 2] = 0u;
                w->cells[4ull * i + 3] = 0u;
                setGasPair(w, i, 0u, 0u);
            }
            CNT(i, G_DEST) = DEST_KEEP;
            continue;
        }
        int drawn = 0;
        if (kindOf(w, t) == KIND_LIQUID) drawn = 1;
        else if (t == E_AIR || (kindOf(w, t) == KIND_GAS && gridZ(w, c) >= 0.0f)) drawn = moveOutOfTheWay(w, i, v);
        if (!drawn) {
            CNT(i, G_DEST) = DEST_KEEP;
            continue;
        }
        uint32_t kinds[RASTER_KINDS] = {NONE, NONE, NONE, NONE}, counts[RASTER_KINDS] = {0, 0, 0, 0};
        for (uint32_t k = start; k < start + n; k++) {
            uint32_t pt = wType(STATE(w->sorted, k, 0)), at = RASTER_KINDS;
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
            if (counts[m] > votes || (MUTANT(10) && counts[m] == votes)) { most = kinds[m]; votes = counts[m]; }
        uint32_t word = 0u, lowest = NONE, solute = 0u, soluteId = NONE;
        int solutes = holdsSolute(most);
        for (uint32_t k = start; k < start + n; k++) {
            uint32_t sx = STATE(w->sorted, k, 0), sz = STATE(w->sorted, k, 2);
            if (wType(sx) == most && (lowest == NONE || (MUTANT(11) ? sz > lowest : sz < lowest))) { word = sx; lowest = sz; }
            if (solutes && wType(sx) == most && wAux(sx) != 0u && (soluteId == NONE || sz < soluteId)) { solute = wAux(sx); soluteId = sz; }
        }
        if (solutes) word = packWord(wType(word), solute, wLife(word), wShade(word), wFlags(word));
        float livesx = 0.0f, livesy = 0.0f, plainX = 0.0f;
        u2 heat = {0, 0}, capOf = {0, 0}, velX = {0, 0}, velY = {0, 0};
        for (uint32_t k = start; k < start + n; k++) {
            uint32_t sx = STATE(w->sorted, k, 0), pt = wType(sx);
            if (pt == most && wAux(sx) == wAux(word)) { livesx += (float)wLife(sx); livesy += 1.0f; }
            else if (pt == most && solutes && wAux(sx) == 0u) { livesy += 1.0f; }
            float hc = fmaxf(WATER_HEATCAP, 1e-6f);
            heat = addExact(heat, hc * fbits(STATE(w->sorted, k, 1)));
            capOf = addExact(capOf, hc);
            velX = addExact(velX, fbits(MOTION(w->sorted, k, 2)));
            plainX += fbits(MOTION(w->sorted, k, 2));
            velY = addExact(velY, fbits(MOTION(w->sorted, k, 3)));
        }
        Rng r;
        seedRng(&r, i, w->P->frame, w->P->seed ^ 0x7a57u);
        uint32_t life = wu32(floorf(livesx / livesy + rnd(&r)));
        life = life < 255u ? life : 255u;
        word = packWord(wType(word), wAux(word), life, wShade(word), wFlags(word));
        float temp = exactValue(heat, EXACT_FINE) / exactValue(capOf, EXACT_FINE);
        v2 vel = V2((MUTANT(18) ? plainX : exactValue(velX, EXACT_FINE)) / (float)n, exactValue(velY, EXACT_FINE) / (float)n);
        w->cells[4ull * i + 0] = wWith(word, FL_DRAWN);
        w->cells[4ull * i + 1] = ubits(temp);
        w->cells[4ull * i + 2] = pack2x16(vel);
        w->cells[4ull * i + 3] = (i << 8) | (n < LIQUID_COUNT_MAX ? n : LIQUID_COUNT_MAX);
        CNT(i, G_TYPE) = (wType(word) | (wAux(word) << 8)) | (life << 16);
    }
}
static void k_unpaint(OWorld *w) {                           /* :4163-4182 */
    FOR_ACTIVE_CELLS(w, c) {
        if (!inWorld(w, c)) continue;
        uint32_t i = cellIndex(w, c);
        uint32_t vx = w->cells[4ull * i];
        if (kindOf(w, wType(vx)) == KIND_LIQUID && CNT(i, G_SORT) == 0u && (CNT(i, G_TAKEN) == 0u || MUTANT(23))) {
            float temp = effTemp(fbits(w->cells[4ull * i + 1]));
            w->cells[4ull * i + 0] = E_AIR;
            w->cells[4ull * i + 1] = ubits(temp);
            w->cells[4ull * i + 2] = 0u;
            w->cells[4ull * i + 3] = 0u;
            setGasPair(w, i, 0u, 0u);
        }
    }
}

/* ------------------------------------------------------------- the step */
/* ------------------------------------------------------- the bodies' passes */
/* bodyStart (:3065-3070): the step's push sums emptied */
static void k_bodyStart(OWorld *w) {
    if (bodiesOn(w)) memset(w->pushes, 0, 4ull * MAX_BODIES * BODY_PUSH_WORDS);
}
/* seepLoads (:3132-3143): the liquid's pressure on free body b's cell c
 * where it rests against what isn't liquid or air */
static void seepLoads(OWorld *w, i2 c, uint32_t b) {
    for (int d = 0; d < 4; d++) {
        i2 q = beside(c, d);
        if (!inWorld(w, q) || rigidAt(w, q) == b || gridZ(w, q) >= 0.0f) continue;
        i2 n = I2(q.x - c.x, q.y - c.y);
        v2 x = V2((float)c.x + 0.5f + 0.5f * (float)n.x, (float)c.y + 0.5f + 0.5f * (float)n.y);
        v3 j = bodyFaceJ(w, q, c, b);
        float s = seepAt(w, c, n, b, 0, x);
        addPush(w, b, PUSH_SURFACE, V3(j.x * s, j.y * s, j.z * s));
    }
}
/* bodyLoads (:3097-3121): the liquid's pressure where each free body's
 * pose puts its surface, past what the cell's middle gives */
static void k_bodyLoads(OWorld *w) {
    if (!bodiesOn(w)) return;
    FOR_ACTIVE_CELLS(w, c) {
        if (!inWorld(w, c)) continue;
        uint32_t own = rigidAt(w, c);
        if (own != NONE) {
            seepLoads(w, c, own);
            continue;
        }
        float g[4];
        gridAt4(w, c, g);
        if (g[2] <= 0.0f) continue;
        v2 mid = V2((float)c.x + 0.5f, (float)c.y + 0.5f);
        for (int d = 0; d < 4; d++) {
            i2 q = beside(c, d);
            uint32_t b = rigidAt(w, q);
            if (b == NONE) continue;
            i2 n = I2(q.x - c.x, q.y - c.y);
            float reach = 1.0f - bodyReach(w, b, I2(-n.x, -n.y));
            v2 surface = V2(mid.x + (float)n.x * reach, mid.y + (float)n.y * reach);
            v3 j = bodyFaceJ(w, c, q, b);
            float dp = potentialAt(w, mid) - potentialAt(w, surface);
            addPush(w, b, PUSH_SURFACE, V3(j.x * g[2] * dp, j.y * g[2] * dp, j.z * g[2] * dp));
        }
    }
}
/* bodyDrag (:3150-3173): the liquid's drag on each free body as the
 * viscosity's sweeps left its faces */
static void k_bodyDrag(OWorld *w) {
    if (!bodiesOn(w)) return;
    FOR_ACTIVE_CELLS(w, c) {
        uint32_t b = rigidAt(w, c);
        if (!inWorld(w, c) || b == NONE) continue;
        v2 pos = bodyV(w, b, BW_POS);
        float rho = bodyDensity(w, b);
        v3 push = V3(0.0f, 0.0f, 0.0f);
        for (int k = 0; k < 4; k++) {
            int isU = k < 2;
            int s = (k & 1) == 1 ? 1 : -1;
            i2 n = isU ? I2(s, 0) : I2(0, s);
            if (gridZ(w, I2(c.x + n.x, c.y + n.y)) <= 0.0f) continue;
            i2 at = isU ? I2(c.x + (k & 1), c.y) : I2(c.x, c.y + (k & 1));
            v2 x = faceMiddle(at, isU);
            int axis = isU ? 0 : 1;
            v2 mv = bodyMoving(w, b, x, 0);
            float du = bodyReach(w, b, n) * rho * (FLD(faceSlot(w, faceIndex(w, at, isU)), 0) - (axis == 0 ? mv.x : mv.y));
            v2 r = V2(x.x - pos.x, x.y - pos.y);
            if (axis == 0) push.x += du;
            else push.y += du;
            push.z += isU ? -r.y * du : r.x * du;
        }
        if (push.x != 0.0f || push.y != 0.0f || push.z != 0.0f) addPush(w, b, PUSH_DRAG, push);
    }
}
/* bodyFaces (:3180-3201): the bodies' faces, as the sweeps left them,
 * moving with the bodies again, the drag and all */
static void k_bodyFaces(OWorld *w) {
    if (!bodiesOn(w)) return;
    FOR_ACTIVE_CELLS(w, c) {
        if (!inWorld(w, c)) continue;
        Face3 fs[4];
        facesOf(w, c, fs);
        for (int j = 0; j < 4; j++) {
            if (fs[j].isU < 0) continue;
            i2 at = I2(fs[j].x, fs[j].y);
            int isU = fs[j].isU == 1;
            i2 a = isU ? I2(at.x - 1, at.y) : I2(at.x, at.y - 1);
            uint32_t ra = rigidAt(w, a), rb = rigidAt(w, at);
            if (ra != NONE || rb != NONE) {
                float ga[4], gb[4];
                gridAt4(w, a, ga);
                gridAt4(w, at, gb);
                setBodyFace(w, faceIndex(w, at, isU), at, isU, ga, gb, ra, rb, 1);
            }
        }
    }
}
/* cgBodyX, cgBodyP, cgBodyF (:2819-2845): the pressure's push on each free
 * body, of the pressure the solve starts from, of an iteration's direction
 * (into the iteration's set by its parity), of the pressure it ends with */
static void k_cgBodyX(OWorld *w) {
    if (!bodiesOn(w)) return;
    FOR_ACTIVE_CELLS(w, c) {
        if (inWorld(w, c) && gridZ(w, c) > 0.0f) addPressurePush(w, c, CG(cellIndex(w, c), 0), PUSH_X);
    }
}
static void k_cgBodyP(OWorld *w) {
    if (!bodiesOn(w) || converged(w)) return;
    FOR_ACTIVE_CELLS(w, c) {
        if (inWorld(w, c) && gridZ(w, c) > 0.0f) addPressurePush(w, c, CG(cellIndex(w, c), 2), PUSH_P0 + w->cycle % 2u);
    }
}
static void k_cgBodyF(OWorld *w) {
    if (!bodiesOn(w)) return;
    FOR_ACTIVE_CELLS(w, c) {
        if (inWorld(w, c) && gridZ(w, c) > 0.0f) addPressurePush(w, c, FLD(2u * cellIndex(w, c), 0), PUSH_F);
    }
}
/* rigidify (:3037-3063): what the liquid's pressure did to each held body,
 * and whether it is wet; summed into its liquid slot's words */
static void k_rigidify(OWorld *w) {
    if (!bodiesOn(w)) return;
    PAR_FOR
    for (uint32_t wid = 0; wid < activeCount(w); wid++) {
        i2 b = activeBlock(w, wid);
        for (int li = 0; li < BLOCK * BLOCK; li++) {
            i2 c = I2(b.x * BLOCK + li % BLOCK, b.y * BLOCK + li / BLOCK);
            uint32_t h = heldAt(w, c);
            if (h == NONE) continue;
            int32_t d[3];
            heldPush(w, c, h, d);
            if ((w->blocks[STATE_SLOT(blockIndex(w, idiv8(c)))] & NEAR_LIQUID) != 0u) at_or(&w->motion[liquidSlot(h, 3u)], LIQUID_WET);
            for (uint32_t k = 0; k < 3u; k++) at_add(&w->motion[liquidSlot(h, k)], (uint32_t)d[k]);
        }
    }
}
/* bodyPush (:3206-3218): what the liquid did to each free body this step */
static void k_bodyPush(OWorld *w) {
    if (!bodiesOn(w)) return;
    for (uint32_t b = 0; b < MAX_BODIES; b++) {
        v3 s = pushChangeFree(w, b, PUSH_SURFACE), dr = pushChangeFree(w, b, PUSH_DRAG), f = pushChangeFree(w, b, PUSH_F);
        v3 ch = V3(s.x + dr.x + f.x, s.y + dr.y + f.y, s.z + dr.z + f.z);
        if (ch.x == 0.0f && ch.y == 0.0f && ch.z == 0.0f) continue;
        w->motion[liquidSlot(b, 0u)] += (uint32_t)wi32(wround(ch.x * RIGID_FIX));
        w->motion[liquidSlot(b, 1u)] += (uint32_t)wi32(wround(ch.y * RIGID_FIX));
        w->motion[liquidSlot(b, 2u)] += (uint32_t)wi32(wround(ch.z * RIGID_FIX));
    }
}

static void relaxPass(OWorld *w, int *shifted, uint32_t times) {
    for (uint32_t k = 0; k < times; k++) {
        k_relaxRims(w);
        k_relax(w, *shifted ? BLOCK / 2 : 0);
        *shifted = !*shifted;
    }
}

static int check_ctl(const OWorld *w, const WigParams *P, const WigStepCtl *ctl) {
    if (P->size[0] != (uint32_t)w->W || P->size[1] != (uint32_t)w->H) return WIG_E_INVALID;
    if ((ctl->surface || ctl->bodies) && ctl->residual_every) return WIG_E_INVALID;
    if (ctl->fluid_cycles > WIG_MAX_CYCLES || ctl->sweep_pairs > 64u || ctl->coarse_sweeps > 4096u || ctl->level_pairs > 64u) return WIG_E_INVALID;
    if (ctl->tol_mode > WIG_TOL_REL_L2 || ctl->residual_every > WIG_MAX_CYCLES) return WIG_E_INVALID;
    return WIG_OK;
}

/* Simulation.encodeFluid (js/sim.js:2466-2612) */
int wor_fluid_step(OWorld *w, const WigParams *P, const WigStepCtl *ctl) {
    int st = check_ctl(w, P, ctl);
    if (st) return st;
    w->P = P;
    w->ctl = ctl;
    w->cycle = 0;
    w->converged_flag = 0;
    if (!ctl->full) {
        k_sync(w, 0);
        k_compact(w);
        k_retire(w);
        return WIG_OK;
    }
    k_sync(w, 1);
    k_syncParticles(w);
    k_syncSeed(w);
    k_compact(w);
    k_retire(w);
    k_clear(w);
    k_gather(w);
    k_sealJoin(w);
    k_sealAnchor(w);
    k_classify(w);
    if (ctl->surface) k_surface(w);
    k_bodyStart(w);
    k_bodyLoads(w);
    k_setFaces(w);
    /* viscosity's sweeps before the projection (js/sim.js:2505-2514) */
    for (int i = 0; ctl->viscous && i < VISCOSITY_ITERATIONS; i++) {
        k_viscosity(w, 0);
        k_viscosity(w, 1);
    }
    k_unpress(w);
    /* the liquid's drag on the bodies, as the sweeps left their faces, and
     * their faces moving with them again (js/sim.js:2515-2520) */
    if (ctl->viscous) {
        k_bodyDrag(w);
        k_bodyFaces(w);
    }
    k_divergence(w);
    int shifted = (int)ctl->shifted_first;
    SolveHead *h = head(w);
    memset(h, 0, sizeof *h);
    k_cgStart(w);
    k_cgJoin(w);
    k_cgAnchor(w);
    k_cgBodyX(w);
    k_cgResidual(w);
    k_cgPrepare(w);
    uint32_t deepest = w->L;
    /* the C ABI's chunked encoding (WigStepCtl.first_chunk): the loop stops
     * at a chunk's end once the solve has converged */
    uint32_t first = ctl->first_chunk == WIG_CHUNK_AUTO ? (w->last_cycles > 1u ? w->last_cycles : 1u) : ctl->first_chunk;
    uint32_t next_check = first ? first : WIG_MAX_CYCLES + 1u;
    for (uint32_t cycle = 0; cycle < ctl->fluid_cycles; cycle++) {
        if (cycle == next_check) {
            if (w->converged_flag) {
                w->last_cycles = h->cycles;
                break;
            }
            next_check = cycle + (ctl->chunk ? ctl->chunk : 1u);
        }
        w->cycle = cycle;
        relaxPass(w, &shifted, ctl->pre_relax);
        uint32_t pairs = ctl->level_pairs ? ctl->level_pairs : 1u;
        for (uint32_t l = 1; l <= deepest; l++) {
            k_restrict(w, l);
            if (l < deepest)
                for (uint32_t k = 0; k < pairs; k++) {
                    k_smooth(w, l, 0);
                    k_smooth(w, l, 1);
                }
        }
        k_coarsest(w, deepest);
        for (uint32_t l = deepest - 1; l >= 1; l--) {
            k_prolong(w, l);
            for (uint32_t k = 0; k < pairs; k++) {
                k_smooth(w, l, 0);
                k_smooth(w, l, 1);
            }
        }
        k_prolongFine(w);
        relaxPass(w, &shifted, ctl->post_relax);
        k_cgRhoBeta(w);
        k_cgDirection(w);
        k_cgBodyP(w);
        k_cgApplyAlpha(w);
        k_cgUpdate(w);
    }
    w->cycle = 0;
    k_cgFinish(w);
    k_cgBodyF(w);
    k_project(w);
    /* what the liquid did to the bodies in it (js/sim.js:2594-2599) */
    if (ctl->bodies_run) {
        k_rigidify(w);
        k_bodyPush(w);
    }
    k_extrapolate(w, 1);
    k_extrapolate(w, 2);
    k_advect(w);
    k_sortCount(w);
    k_scanBlocks(w);
    k_scanTotals(w);
    k_scanFinish(w);
    k_scatter(w);
    k_rasterClaims(w);
    k_raster(w);
    k_unpaint(w);
    return w->parts[0] > w->cap ? WIG_WARN_OVERFLOW : WIG_OK;
}

/* ------------------------------------------------------------------- API */
OWorld *wor_create(uint32_t width, uint32_t height, uint32_t coarse_levels) {
    if (width < 2u || height < 2u || width > 16384u || height > 16384u || coarse_levels > MG_LEVELS) return NULL;
    OWorld *w = (OWorld *)calloc(1, sizeof(OWorld));
    if (!w) return NULL;
    layout(w, width, height, coarse_levels);
    w->cells = (uint32_t *)calloc(1, w->bytes[WIG_PART_CELLS]);
    w->sorted = (uint32_t *)calloc(1, w->bytes[WIG_PART_PARTICLES0]);
    w->parts = (uint32_t *)calloc(1, w->bytes[WIG_PART_PARTICLES1]);
    w->counters = (uint32_t *)calloc(1, w->bytes[WIG_PART_COUNTERS]);
    w->field = (float *)calloc(1, w->bytes[WIG_PART_FIELD]);
    w->blocks = (uint32_t *)calloc(1, w->bytes[WIG_PART_BLOCKS]);
    w->cg = (float *)calloc(1, w->bytes[WIG_PART_FLUIDCG]);
    w->solve = (uint8_t *)calloc(1, w->bytes[WIG_PART_SOLVE]);
    w->bvec = (float *)calloc(w->C, sizeof(float));
    w->air = (float *)calloc(1, w->bytes[WIG_PART_AIR]);
    w->source = (uint32_t *)calloc(1, w->bytes[WIG_PART_AIR_SOURCE]);
    w->occ = (uint32_t *)calloc(1, w->bytes[WIG_PART_OCCUPANTS]);
    w->grav = (float *)calloc(1, w->bytes[WIG_PART_GRAVITY]);
    w->motion = (uint32_t *)calloc(1, w->bytes[WIG_PART_MOTION]);
    w->swept = (uint32_t *)calloc(1, w->bytes[WIG_PART_SWEPT]);
    w->pushes = (uint32_t *)calloc(1, w->bytes[WIG_PART_PUSHES]);
    w->dot_mode = 1;
    if (!w->cells || !w->sorted || !w->parts || !w->counters || !w->field || !w->blocks || !w->cg || !w->solve || !w->bvec
        || !w->air || !w->source || !w->occ || !w->grav || !w->motion || !w->swept || !w->pushes)
        return NULL;
    return w;
}
void wor_destroy(OWorld *w) {
    if (!w) return;
    free(w->cells); free(w->sorted); free(w->parts); free(w->counters);
    free(w->field); free(w->blocks); free(w->cg); free(w->solve); free(w->cap_b); free(w->bvec);
    free(w->air); free(w->source); free(w->occ); free(w->grav);
    free(w->motion); free(w->swept); free(w->pushes);
    free(w);
}
void *wor_part(OWorld *w, int part, uint64_t *bytes) {
    if (part < 0 || part >= WIG_PART_COUNT) return NULL;
    if (bytes) *bytes = w->bytes[part];
    void *p[WIG_PART_COUNT] = {w->cells, w->sorted, w->parts, w->counters, w->field, w->blocks, w->cg, w->solve,
                               w->air, w->source, w->occ, w->grav, w->motion, w->swept, w->pushes};
    return p[part];
}
void wor_set_kinds(OWorld *w, const uint8_t kinds[256]) { memcpy(w->kind, kinds, 256); }
void wor_set_element_props(OWorld *w, const WigElementProps props[256]) { memcpy(w->props, props, sizeof w->props); }
void wor_set_dot_mode(OWorld *w, int mode) { w->dot_mode = mode ? 1 : 0; }
int wor_layout(const OWorld *w, WigLayoutInfo *out) {
    memset(out, 0, sizeof *out);
    out->width = (uint32_t)w->W;
    out->height = (uint32_t)w->H;
    out->cells = w->C;
    out->faces = w->F;
    out->blocks = w->B;
    out->blocks_wide = w->bw;
    out->capacity = w->cap;
    out->levels = w->L;
    for (uint32_t l = 0; l <= w->L; l++) { out->level_width[l] = w->lw[l]; out->level_height[l] = w->lh[l]; }
    for (int p = 0; p < WIG_PART_COUNT; p++) out->part_bytes[p] = w->bytes[p];
    return WIG_OK;
}
/* Keeps each step's right-hand side b (= -divergence at the fluid cells,
 * 0 elsewhere) for the FP64 CG oracle below. */
void wor_capture(OWorld *w, int on) {
    w->capture = on;
    if (on && !w->cap_b) w->cap_b = (float *)calloc(w->C, sizeof(float));
    if (on && w->cap_b) memset(w->cap_b, 0, sizeof(float) * w->C);
}
const float *wor_captured_b(OWorld *w) { return w->cap_b; }

/* ------------------------------------------------------ FP64 CG oracle */
/*
 * Plain conjugate gradients in FP64 on the matrix the step's solve used,
 * rebuilt from the world's state after the step: A_ij = -beta_ij between
 * fluid cells, A_ii = sum of beta over non-obstacle neighbours plus PIN at
 * the lowest-index cell of each fluid region that meets no air cell, with
 * beta = fineBeta (fluid.wgsl:396-405,2594-2613, 2557-2592).  The regions
 * are found here by a breadth-first search, independently of the solve's
 * union-find.  Solves A x = b to ||r|| <= tol ||b||; returns iterations,
 * -1 on failure.
 */
int wor_cg64(OWorld *w, const float *b, double *x, double tol, int maxit, double *relres) {
    uint32_t C = w->C;
    int32_t *id = (int32_t *)malloc(sizeof(int32_t) * C);
    uint8_t *pin = (uint8_t *)calloc(C, 1);
    uint32_t *queue = (uint32_t *)malloc(sizeof(uint32_t) * C);
    double *r = (double *)calloc(C, sizeof(double)), *p = (double *)calloc(C, sizeof(double)), *q = (double *)calloc(C, sizeof(double));
    if (!id || !pin || !queue || !r || !p || !q) return -1;
    for (uint32_t i = 0; i < C; i++) id[i] = -1;
    for (uint32_t s = 0; s < C; s++) {
        if (id[s] >= 0 || FLD(2u * s, 2) <= 0.0f) continue;
        uint32_t head_ = 0, tail = 0;
        queue[tail++] = s;
        id[s] = (int32_t)s;
        int open = 0;
        while (head_ < tail) {
            uint32_t i = queue[head_++];
            i2 c = I2((int32_t)(i % (uint32_t)w->W), (int32_t)(i / (uint32_t)w->W));
            for (int d = 0; d < 4; d++) {
                i2 n = beside(c, d);
                if (!inWorld(w, n)) continue;
                uint32_t j = cellIndex(w, n);
                float z = FLD(2u * j, 2);
                if (z == 0.0f) open = 1;
                if (z > 0.0f && id[j] < 0) { id[j] = (int32_t)s; queue[tail++] = j; }
            }
        }
        if (!open) pin[s] = 1;     /* s is the region's lowest index: found in index order */
    }
    /* y = A v on fluid cells */
#define APPLY(v, y)                                                                      \
    for (uint32_t i = 0; i < C; i++) {                                                   \
        y[i] = 0.0;                                                                      \
        float rho = FLD(2u * i, 2);                                                      \
        if (rho <= 0.0f) continue;                                                       \
        i2 c = I2((int32_t)(i % (uint32_t)w->W), (int32_t)(i / (uint32_t)w->W));         \
        double diag = pin[i] ? (double)PIN : 0.0, off = 0.0;                             \
        for (int d = 0; d < 4; d++) {                                                    \
            i2 n = beside(c, d);                                                         \
            if (!inWorld(w, n)) continue;                                                \
            float rn = FLD(2u * cellIndex(w, n), 2);                                     \
            if (rn < 0.0f) continue;                                                     \
            double beta = (double)fineBeta(rho, rn, densityAt(w, c), densityAt(w, n));   \
            diag += beta;                                                                \
            if (rn > 0.0f) off += beta * v[cellIndex(w, n)];                             \
        }                                                                                \
        y[i] = diag * v[i] - off;                                                        \
    }
    double bb = 0.0;
    for (uint32_t i = 0; i < C; i++) {
        x[i] = 0.0;
        r[i] = FLD(2u * i, 2) > 0.0f ? (double)b[i] : 0.0;
        p[i] = r[i];
        bb += r[i] * r[i];
    }
    double rr = bb;
    int it = 0;
    while (it < maxit && rr > tol * tol * bb && bb > 0.0) {
        APPLY(p, q);
        double pq = 0.0;
        for (uint32_t i = 0; i < C; i++) pq += p[i] * q[i];
        if (pq <= 0.0) break;
        double alpha = rr / pq, rn = 0.0;
        for (uint32_t i = 0; i < C; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            rn += r[i] * r[i];
        }
        double beta = rn / rr;
        rr = rn;
        for (uint32_t i = 0; i < C; i++) p[i] = r[i] + beta * p[i];
        it++;
    }
    /* the true residual */
    APPLY(x, q);
    double tr = 0.0;
    for (uint32_t i = 0; i < C; i++) {
        if (FLD(2u * i, 2) <= 0.0f) continue;
        double e = (double)b[i] - q[i];
        tr += e * e;
    }
    if (relres) *relres = bb > 0.0 ? sqrt(tr / bb) : 0.0;
#undef APPLY
    free(id); free(pin); free(queue); free(r); free(p); free(q);
    return it;
}

/* ||b - A x|| / ||b|| for a given x (float), with the same A. */
double wor_relres(OWorld *w, const float *b, const float *xf) {
    uint32_t C = w->C;
    double *x = (double *)malloc(sizeof(double) * C);
    if (!x) return -1.0;
    for (uint32_t i = 0; i < C; i++) x[i] = xf[i];
    /* reuse cg64's pinning by running zero iterations on a copy is awkward:
     * rebuild inline */
    int32_t *id = (int32_t *)malloc(sizeof(int32_t) * C);
    uint8_t *pin = (uint8_t *)calloc(C, 1);
    uint32_t *queue = (uint32_t *)malloc(sizeof(uint32_t) * C);
    for (uint32_t i = 0; i < C; i++) id[i] = -1;
    for (uint32_t s = 0; s < C; s++) {
        if (id[s] >= 0 || FLD(2u * s, 2) <= 0.0f) continue;
        uint32_t hd = 0, tl = 0;
        queue[tl++] = s;
        id[s] = (int32_t)s;
        int open = 0;
        while (hd < tl) {
            uint32_t i = queue[hd++];
            i2 c = I2((int32_t)(i % (uint32_t)w->W), (int32_t)(i / (uint32_t)w->W));
            for (int d = 0; d < 4; d++) {
                i2 n = beside(c, d);
                if (!inWorld(w, n)) continue;
                uint32_t j = cellIndex(w, n);
                float z = FLD(2u * j, 2);
                if (z == 0.0f) open = 1;
                if (z > 0.0f && id[j] < 0) { id[j] = (int32_t)s; queue[tl++] = j; }
            }
        }
        if (!open) pin[s] = 1;
    }
    double num = 0.0, den = 0.0;
    for (uint32_t i = 0; i < C; i++) {
        float rho = FLD(2u * i, 2);
        if (rho <= 0.0f) continue;
        i2 c = I2((int32_t)(i % (uint32_t)w->W), (int32_t)(i / (uint32_t)w->W));
        double diag = pin[i] ? (double)PIN : 0.0, off = 0.0;
        for (int d = 0; d < 4; d++) {
            i2 n = beside(c, d);
            if (!inWorld(w, n)) continue;
            float rn = FLD(2u * cellIndex(w, n), 2);
            if (rn < 0.0f) continue;
            double beta = (double)fineBeta(rho, rn, densityAt(w, c), densityAt(w, n));
            diag += beta;
            if (rn > 0.0f) off += beta * x[cellIndex(w, n)];
        }
        double e = (double)b[i] - (diag * x[i] - off);
        num += e * e;
        den += (double)b[i] * b[i];
    }
    free(x); free(id); free(pin); free(queue);
    return den > 0.0 ? sqrt(num / den) : sqrt(num);
}

/* Known-answer hooks for the integer functions the tests pin. */
uint32_t wor_pcg(uint32_t v) { return pcg(v); }
uint32_t wor_seed_id(OWorld *w, int32_t x, int32_t y, uint32_t frame, uint32_t j) {
    WigParams P;
    memset(&P, 0, sizeof P);
    P.frame = frame;
    w->P = &P;
    uint32_t r = seedId(w, I2(x, y), j);
    w->P = NULL;
    return r;
}
void wor_seed_pos(int32_t x, int32_t y, uint32_t i, uint32_t frame, uint32_t seed, uint32_t n, float *out) {
    Rng r;
    seedRng(&r, i, frame, seed ^ 0x5eedu);
    for (uint32_t j = 0; j < n; j++) {
        v2 p = seedPos(&r, I2(x, y), j, n);
        out[2 * j] = p.x;
        out[2 * j + 1] = p.y;
    }
}
uint16_t wor_f16(float f) { return f16_of(f); }
float wor_f32(uint16_t h) { return f32_of(h); }
void wor_exact_parts(float x, float scale, uint32_t *out) { u2 p = exactParts(x, scale); out[0] = p.lo; out[1] = p.hi; }
float wor_exact_value(uint32_t lo, uint32_t hi, float scale) { u2 w = {lo, hi}; return exactValue(w, scale); }
float wor_fine_beta(float a, float b, float da, float db) { return fineBeta(a, b, da, db); }

/* --------------------------------------------- one pass (test hooks)
 * The tests that pin a pass by itself build a state by hand and run that
 * pass alone on it, with the step's params and controls (iteration 0 of the
 * solve, not converged).  WOR_PASS_* ids; `arg` is the pass's level (restrict,
 * smooth, coarsest, prolong), layer (extrapolate), or whether sync takes claims;
 * smooth takes the colour in arg's bit 8.  The sort runs sortCount through
 * scatter, the seal sealJoin then sealAnchor. */
enum {
    WOR_PASS_SYNC, WOR_PASS_SYNC_PARTICLES, WOR_PASS_SYNC_SEED, WOR_PASS_COMPACT, WOR_PASS_RETIRE, WOR_PASS_CLEAR,
    WOR_PASS_GATHER, WOR_PASS_SEAL, WOR_PASS_CLASSIFY, WOR_PASS_SET_FACES, WOR_PASS_UNPRESS, WOR_PASS_DIVERGENCE,
    WOR_PASS_RESTRICT, WOR_PASS_SMOOTH, WOR_PASS_COARSEST, WOR_PASS_PROLONG, WOR_PASS_PROLONG_FINE, WOR_PASS_PROJECT,
    WOR_PASS_EXTRAPOLATE, WOR_PASS_ADVECT, WOR_PASS_SORT, WOR_PASS_RASTER_CLAIMS, WOR_PASS_RASTER, WOR_PASS_UNPAINT,
    WOR_PASS_RELAX, WOR_PASS_VISCOSITY, WOR_PASS_SURFACE, WOR_PASSES
};
int wor_run(OWorld *w, int pass, int arg, const WigParams *P, const WigStepCtl *ctl) {
    if (pass < 0 || pass >= WOR_PASSES) return WIG_E_INVALID;
    w->P = P;
    w->ctl = ctl;
    w->cycle = 0;
    w->converged_flag = 0;
    const uint32_t l = (uint32_t)arg & 0xffu;
    if ((pass == WOR_PASS_RESTRICT || pass == WOR_PASS_SMOOTH || pass == WOR_PASS_COARSEST || pass == WOR_PASS_PROLONG) &&
        (l < 1u || l > w->L || (pass == WOR_PASS_RESTRICT && l < 1u) || (pass == WOR_PASS_PROLONG && l >= w->L)))
        return WIG_E_INVALID;
    switch (pass) {
    case WOR_PASS_SYNC: k_sync(w, arg); break;
    case WOR_PASS_SYNC_PARTICLES: k_syncParticles(w); break;
    case WOR_PASS_SYNC_SEED: k_syncSeed(w); break;
    case WOR_PASS_COMPACT: k_compact(w); break;
    case WOR_PASS_RETIRE: k_retire(w); break;
    case WOR_PASS_CLEAR: k_clear(w); break;
    case WOR_PASS_GATHER: k_gather(w); break;
    case WOR_PASS_SEAL: k_sealJoin(w); k_sealAnchor(w); break;
    case WOR_PASS_CLASSIFY: k_classify(w); break;
    case WOR_PASS_SET_FACES: k_setFaces(w); break;
    case WOR_PASS_UNPRESS: k_unpress(w); break;
    case WOR_PASS_DIVERGENCE: k_divergence(w); break;
    case WOR_PASS_RESTRICT: k_restrict(w, l); break;
    case WOR_PASS_SMOOTH: k_smooth(w, l, (arg >> 8) & 1); break;
    case WOR_PASS_COARSEST: k_coarsest(w, l); break;
    case WOR_PASS_PROLONG: k_prolong(w, l); break;
    case WOR_PASS_PROLONG_FINE: k_prolongFine(w); break;
    case WOR_PASS_PROJECT: k_project(w); break;
    case WOR_PASS_EXTRAPOLATE: k_extrapolate(w, arg == 2 ? 2 : 1); break;
    case WOR_PASS_ADVECT: k_advect(w); break;
    case WOR_PASS_SORT: k_sortCount(w); k_scanBlocks(w); k_scanTotals(w); k_scanFinish(w); k_scatter(w); break;
    case WOR_PASS_RASTER_CLAIMS: k_rasterClaims(w); break;
    case WOR_PASS_RASTER: k_raster(w); break;
    case WOR_PASS_UNPAINT: k_unpaint(w); break;
    case WOR_PASS_RELAX: k_relax(w, arg ? BLOCK / 2 : 0); break;
    case WOR_PASS_VISCOSITY: k_viscosity(w, arg & 1); break;
    case WOR_PASS_SURFACE: k_surface(w); break;
    }
    return WIG_OK;
}
/* Where coarse level l (1 .. levels) starts in the field, in vec4s. */
uint64_t wor_level_start(const OWorld *w, uint32_t l) { return l >= 1u && l <= w->L ? w->lstart[l] : 0u; }

/* The bodies' helpers one at a time, for the pins in tests/ (the step's
 * Params and controls as a pass sees them):
 *   0 bodyFaceJ(c (i0, i1), q (i2, i3), body i4)          -> f[0..2]
 *   1 couplingAt(c (i0, i1), set i2)                      -> f[0]
 *   2 addPressurePush(c (i0, i1), p f0, set i2)
 *   3 rigidAt(c (i0, i1)), 4 heldAt(c (i0, i1))           -> u[0]
 *   5 the swept body of cell i0, i1 steps back, seam i2, weighed i3 -> u[0]
 *   6 giveWay(body i0, g (f0, f1, f2))                    -> f[0..2]
 *   7 bodyMoving(body i0, x (f0, f1), drag i1)            -> f[0..1]
 *   8 potentialAt((f0, f1))                               -> f[0]
 *   9 heldPush(c (i0, i1), body i2)                       -> u[0..2] (i32)
 *  10 rigidBodyOf(cell i0's word)                         -> u[0]
 *  11 seepAt(c (i0, i1), n (i2, i3), body i4, held i5, x (f0, f1)) -> f[0]
 *  12 pushOf(body i0, set i1)                             -> f[0..2] */
int wor_body_probe(OWorld *w, const WigParams *P, const WigStepCtl *ctl, int op, const int32_t *ia, const float *fa,
                   float *fo, uint32_t *uo) {
    w->P = P;
    w->ctl = ctl;
    i2 c = I2(ia[0], ia[1]);
    switch (op) {
    case 0: { v3 j = bodyFaceJ(w, c, I2(ia[2], ia[3]), (uint32_t)ia[4]); fo[0] = j.x; fo[1] = j.y; fo[2] = j.z; break; }
    case 1: fo[0] = couplingAt(w, c, (uint32_t)ia[2]); break;
    case 2: addPressurePush(w, c, fa[0], (uint32_t)ia[2]); break;
    case 3: uo[0] = rigidAt(w, c); break;
    case 4: uo[0] = heldAt(w, c); break;
    case 5: uo[0] = sweptBody(w, (uint32_t)ia[0], (uint32_t)ia[1], (uint32_t)ia[2], ia[3]); break;
    case 6: { v3 g = giveWay(w, (uint32_t)ia[0], V3(fa[0], fa[1], fa[2])); fo[0] = g.x; fo[1] = g.y; fo[2] = g.z; break; }
    case 7: { v2 m = bodyMoving(w, (uint32_t)ia[0], V2(fa[0], fa[1]), ia[1]); fo[0] = m.x; fo[1] = m.y; break; }
    case 8: fo[0] = potentialAt(w, V2(fa[0], fa[1])); break;
    case 9: { int32_t o[3]; heldPush(w, c, (uint32_t)ia[2], o); for (int k = 0; k < 3; k++) uo[k] = (uint32_t)o[k]; break; }
    case 10: uo[0] = rigidBodyOf(w, &w->cells[4ull * (uint32_t)ia[0]]); break;
    case 11: fo[0] = seepAt(w, c, I2(ia[2], ia[3]), (uint32_t)ia[4], ia[5], V2(fa[0], fa[1])); break;
    case 12: { v3 g = pushOf(w, (uint32_t)ia[0], (uint32_t)ia[1]); fo[0] = g.x; fo[1] = g.y; fo[2] = g.z; break; }
    default: return WIG_E_INVALID;
    }
    return WIG_OK;
}
