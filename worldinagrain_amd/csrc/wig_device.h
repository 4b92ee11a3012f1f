// Device-side definitions of the liquid step for gfx950: the world's view
// (sizes and buffer pointers, passed by value to every kernel), WGSL's
// arithmetic (round to even, saturating conversions, mix, f16 packing), and
// the reference's index math (shaders/fluid.wgsl:246-513, 2228-2248).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wig.h"

#define WIG_BLOCK 8
#define LIQUID_CELL_WORDS 11u
#define LIQUID_PER_CELL 4u
#define FLUID_STATS 9u
#define MG_COARSEST 512
#define MG_LEVELS 8
#define CF_FLUID 1u
#define CF_PINNED 2u
#define CF_BODY 4u      // a fluid cell beside a free body's (the solve couples it: couplingAt)
#define LDS_LEVEL_FLOATS 7u   // k_coarseLDS: p, rhs (2 each), rho, two face betas a cell
#define WNONE 0xffffffffu
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
#define FIXF 65536.0f
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
#define RELAX_NT 1       // tiles a wavefront of k_relax works at once (four wavefronts a workgroup)
#define PINF 1.0f
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
#define DEST_KEEP 0xffffffffu
#define DEST_ASIDE 0x80000000u
#define DEST_CELL 0x40000000u
#define DEST_SPAN 0x3fffffffu
#define LIQUID_COUNT_MAX 255u
#define FL_DRAWN 4u
#define KIND_SOLID 1u
#define KIND_POWDER 2u
#define KIND_LIQUID 3u
#define KIND_GAS 4u
#define E_AIR 0u
#define E_WATER 3u
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
#define ATTAINABLE2 8.8817841970012523e-16   // (2^-25)^2: residual replacement's stop at twice an f32 iterate's rounding floor
// the air (js/shared.js:58,137-175,238,289; common.wgsl:287-288;
// js/elements.js:424,449-457,912-914: the packed table's f32 values)
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
// the bodies (js/shared.js:294-333; common.wgsl:353,377-387; fluid.wgsl:1304-1318,2963)
#define MAX_BODIES WIG_MAX_BODIES
#define BODY_TABLE_WORDS 116u
#define BODY_CONTACTS 16384u
#define BODY_CONTACT_WORDS 12u
#define BODY_LIQUID_WORDS 4u
#define BODY_PUSH_WORDS WIG_BODY_PUSH_WORDS
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

// The FP64 state of a solve: its header, then three doubles per block
// (two running sums of a pass and the squared residual).
struct SolveHead {
    double rz, alpha, beta, pq, rr, bb;
    uint32_t cycles, converged;
    float vscale;       // 16-bit V-cycle storage: the power of two its pressure rhs is scaled by
    uint32_t pad;
};

// A world's own device words that outlive a step: how many iterations its
// last solve stopped after, and this solve's first convergence checkpoint
// and the iteration it stops before (WIG_CHUNK_AUTO and chunk, include/wig.h)
#define WS_LAST 0
#define WS_FIRST 1
#define WS_STOP 2
#define WS_CYCLE 3      // the solve's iteration: set by cgResidual, stepped by cgCheck
#define WS_WORDS 16
#define NO_STOP 0xffffffffu

// Everything a kernel needs to know of the world and the step.
struct Dev {
    int32_t W, H;
    uint32_t C, F, B, bw, bh, cap, L;
    uint32_t lw[MG_LEVELS + 1], lh[MG_LEVELS + 1];
    uint64_t lstart[MG_LEVELS + 2];
    uint32_t *cells;
    uint32_t *sorted;
    uint32_t *parts;
    uint32_t *counters;
    float *field;
    uint32_t *blocks;
    float *cg;
    SolveHead *head;
    double *bpart;
    const uint8_t *kind;
    const WigElementProps *eprops;   // the elements' liquid columns (wig_set_element_props)
    uint32_t *scratch;      // per-block flags and offsets (compact)
    unsigned long long *totals;  // steps, PCG iterations, particles, active cells (run totals)
    uint32_t *wstate;       // WS_* words
    // The solve's fine level, packed once a solve (cgPrepare) for the cells
    // of the blocks worked on and the ring of cells beside them: each cell's
    // four face betas in beside()'s order (fineBeta; the sign bit set where
    // the neighbour is fluid), its fluid and pinned bits, the V-cycle's
    // iterate (p, drift) in two buffers (a relaxation reads one and writes
    // the other, so its tiles' rims stay as they were), and the rhs (-r, drift
    // rhs).  The reference's words for these are written back at cgFinish.
    float4 *coef;
    uint8_t *cflag;
    float2 *pa, *pb, *r2;
    // and the CG vectors (fluidCg's x, r, p, Ap), p in two buffers: an
    // iteration's cgApply reads the last p and writes the new one
    float *cx, *cr, *cq, *cp0, *cp1;
    // The V-cycle's storage (WIG_STORE_*): with 16-bit storage its fine
    // iterate and rhs (two halves a cell, as pa, pb, r2) and its coarse
    // levels (p, rhs as four halves a cell, rho as a float) live here
    // instead of the float arrays and the field's level records.
    uint32_t store;
    // The FP64 check of a solve (wig_cg64): the solve's rhs as cgResidual
    // saw it (when capturing), and FP64 CG vectors and per-block partials
    float *capb;
    double *x64, *r64, *p64, *q64, *part64, *head64;
    uint32_t *ha, *hb, *hr2;
    uint16_t *vlev[MG_LEVELS + 1];
    float *vrho[MG_LEVELS + 1];
    // a tiled level's tiles: whether any of a tile's own cells holds fluid,
    // taken by k_levelDown in each solve's first iteration (tile by row)
    uint8_t *tflag[MG_LEVELS + 1];
    // a tiled level's correction as k_levelDown leaves it, for k_levelUp
    // (8 B a cell: two floats, or two 16-bit values in the first word).
    // The level's own records take k_levelUp's result; were the down pass's
    // kept there too, a tile of either pass could read into its halo a value
    // a neighbouring tile of the same launch had already overwritten.
    uint2 *pdown[MG_LEVELS + 1];
    // a tiled level's face betas (faceBeta, of its right and lower faces; 0
    // where the stencil skips), taken by k_levelDown in each solve's first
    // iteration: level rho does not change within a solve
    float2 *vbeta[MG_LEVELS + 1];
    uint32_t lds_level;     // the first coarse level solved in one workgroup's LDS
    uint32_t wave_cells;    // and below it, levels of at most this many cells by its first wavefront
    uint32_t tile[MG_LEVELS + 1];   // k_levelDown/k_levelUp's tile side, per level below lds_level
    // the step's: the Params words the liquid reads, and its controls
    uint32_t frame, seed;
    float earthG;
    uint32_t full, fluid_cycles, shifted_first, first_chunk, chunk;
    uint32_t sweep_pairs, coarse_sweeps, tol_mode, level_pairs;
    float tol;
    // residual replacement (WigStepCtl.residual_every; 0 off) and the
    // solve's b (packed), which it reads
    uint32_t replace_every;
    float *cb;
    // the air (WigStepCtl.air): air[0] and airSource, the wind's drag
    // coefficient, and the step's dt and dx (the pressure's push)
    const float4 *air;
    const uint32_t *occ;         // occupants (WIG_PART_OCCUPANTS), read when occ_on
    const float4 *grav;          // the gravity pass's field (WIG_PART_GRAVITY), read when grav_on
    uint32_t *source;
    uint32_t air_on, occ_on, grav_on, visc_on, surf_on;
    // the bodies (WigStepCtl.bodies): the bodies' table and liquid slots, the
    // swept words, the step's push sums; read when body_on
    uint32_t *motion;
    const uint32_t *swept;
    uint32_t *pushes;
    uint32_t body_on, bodies_run;
    float surface_drag, dt, dx;
};

// A launch over a batch of worlds.  Its workgroups are split among the
// worlds in contiguous runs, one run a world, laid out per launch shape:
// `map` gives each workgroup's world and `first` each world's first
// workgroup (and, at nworlds, the total).  A kernel takes its world's Dev
// from `tab` and its workgroup's index and count within the world's run.
struct WorldGrid {
    const Dev *__restrict__ tab;
    const uint32_t *__restrict__ map;
    const uint32_t *__restrict__ first;
    uint32_t solo;      // one world: tab[0], the whole grid its run (no map to read first)
};
// (d is the workgroup's own copy, read once at the kernel's start: the
// kernels' stores could otherwise be taken to change it, and each use of a
// field would read it again from memory)
#define WORLD(g)                                                      \
    const uint32_t wi_ = (g).solo ? 0u : (g).map[blockIdx.x];         \
    const Dev d = (g).tab[wi_];                                       \
    const uint32_t bid = (g).solo ? blockIdx.x : blockIdx.x - (g).first[wi_];          \
    const uint32_t nbid = (g).solo ? gridDim.x : (g).first[wi_ + 1u] - (g).first[wi_]; \
    (void)nbid; (void)bid
// The level kernels' view: level l of the world as level 1 of the copy and
// level l + 1 as its level 2 (a copy indexed by a level known only at run
// time would live in scratch; constant indices keep it in registers), and
// level l's tile side.
#define WORLD_LEVEL(g, l)                                             \
    const uint32_t wi_ = (g).solo ? 0u : (g).map[blockIdx.x];         \
    const Dev &dg_ = (g).tab[wi_];                                    \
    Dev d = dg_;                                                      \
    d.lw[1] = dg_.lw[l];                                              \
    d.lh[1] = dg_.lh[l];                                              \
    d.lstart[1] = dg_.lstart[l];                                      \
    d.vlev[1] = dg_.vlev[l];                                          \
    d.vrho[1] = dg_.vrho[l];                                          \
    d.lw[2] = dg_.lw[(l) + 1u];                                       \
    d.lh[2] = dg_.lh[(l) + 1u];                                       \
    d.lstart[2] = dg_.lstart[(l) + 1u];                               \
    d.vlev[2] = dg_.vlev[(l) + 1u];                                   \
    d.vrho[2] = dg_.vrho[(l) + 1u];                                   \
    d.tflag[1] = dg_.tflag[l];                                        \
    d.pdown[1] = dg_.pdown[l];                                        \
    d.vbeta[1] = dg_.vbeta[l];                                        \
    const uint32_t tile_ = dg_.tile[l];                               \
    const uint32_t bid = (g).solo ? blockIdx.x : blockIdx.x - (g).first[wi_];          \
    const uint32_t nbid = (g).solo ? gridDim.x : (g).first[wi_ + 1u] - (g).first[wi_]; \
    (void)nbid; (void)bid
// (the Dev in memory, for kernels that index its per-level arrays by a
// level known only at run time throughout)
#define WORLD_REF(g)                                                  \
    const uint32_t wi_ = (g).solo ? 0u : (g).map[blockIdx.x];         \
    const Dev &d = (g).tab[wi_];                                      \
    const uint32_t bid = (g).solo ? blockIdx.x : blockIdx.x - (g).first[wi_];          \
    const uint32_t nbid = (g).solo ? gridDim.x : (g).first[wi_ + 1u] - (g).first[wi_]; \
    (void)nbid; (void)bid

struct i2 { int32_t x, y; };

// 16-bit storage of V-cycle values: IEEE half (round to nearest even, the
// hardware conversion) or bfloat16 (the float's top half, rounded to nearest
// even); arithmetic stays float.
#define STORE_F32 0
#define STORE_F16 1
#define STORE_BF16 2
template <int S> __device__ __forceinline__ uint16_t enc16(float v) {
    if (S == STORE_F16) {
        _Float16 h = (_Float16)v;
        return __builtin_bit_cast(uint16_t, h);
    }
    uint32_t u = __float_as_uint(v);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);   // NaN stays NaN
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
template <int S> __device__ __forceinline__ float dec16(uint16_t v) {
    if (S == STORE_F16) return (float)__builtin_bit_cast(_Float16, v);
    return __uint_as_float((uint32_t)v << 16);
}
struct v2 { float x, y; };
__device__ __forceinline__ i2 I2(int32_t x, int32_t y) { i2 r; r.x = x; r.y = y; return r; }
__device__ __forceinline__ v2 V2(float x, float y) { v2 r; r.x = x; r.y = y; return r; }

// ---------------------------------------------------------------- WGSL arithmetic
__device__ __forceinline__ float wround(float x) { return __builtin_rintf(x); }
__device__ __forceinline__ float wclamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float wmix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float wlength(v2 v) { return sqrtf(v.x * v.x + v.y * v.y); }
__device__ __forceinline__ int32_t wi32(float x) {
    if (x != x) return 0;
    if (x <= -2147483648.0f) return INT32_MIN;
    if (x >= 2147483648.0f) return INT32_MAX;
    return (int32_t)x;
}
__device__ __forceinline__ uint32_t wu32(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)x;
}
__device__ __forceinline__ float fbits(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t ubits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ uint32_t pack2x16(v2 v) {
    _Float16 a = (_Float16)v.x, b = (_Float16)v.y;
    return (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}
__device__ __forceinline__ v2 unpack2x16(uint32_t u) {
    _Float16 a = __builtin_bit_cast(_Float16, (uint16_t)(u & 0xffffu));
    _Float16 b = __builtin_bit_cast(_Float16, (uint16_t)(u >> 16));
    return V2((float)a, (float)b);
}

// cell words (js/shared.js:79,499-531)
__device__ __forceinline__ uint32_t wType(uint32_t x) { return x & 0xffu; }
__device__ __forceinline__ uint32_t wAux(uint32_t x) { return (x >> 8) & 0xffu; }
__device__ __forceinline__ uint32_t wLife(uint32_t x) { return (x >> 16) & 0xffu; }
__device__ __forceinline__ uint32_t wShade(uint32_t x) { return (x >> 24) & 0xfu; }
__device__ __forceinline__ uint32_t wFlags(uint32_t x) { return x >> 28; }
__device__ __forceinline__ uint32_t packWord(uint32_t t, uint32_t aux, uint32_t life, uint32_t shade, uint32_t flags) {
    return (t & 0xffu) | ((aux & 0xffu) << 8) | ((life & 0xffu) << 16) | ((shade & 0xfu) << 24) | ((flags & 0xfu) << 28);
}
__device__ __forceinline__ uint32_t wWith(uint32_t x, uint32_t f) { return x | (f << 28); }
__device__ __forceinline__ uint32_t wWithout(uint32_t x, uint32_t f) { return x & ~(f << 28); }
__device__ __forceinline__ bool wHas(uint32_t x, uint32_t f) { return (wFlags(x) & f) != 0u; }

// the PCG hash (common.wgsl:1415-1432)
__device__ __forceinline__ uint32_t pcg(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28) + 4u)) ^ s) * 277803737u;
    return (w >> 22) ^ w;
}
struct Rng { uint32_t s; };
__device__ __forceinline__ void seedRng(Rng &r, uint32_t a, uint32_t b, uint32_t c) { r.s = pcg(a ^ pcg(b ^ pcg(c))); }
__device__ __forceinline__ uint32_t rndU(Rng &r) { r.s = pcg(r.s); return r.s; }
__device__ __forceinline__ float rnd(Rng &r) { return (float)(rndU(r) >> 8) * (1.0f / 16777216.0f); }

// exact sums (common.wgsl:1379-1411)
struct u2 { uint32_t lo, hi; };
__device__ __forceinline__ u2 exactParts(float x, float scale) {
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
__device__ __forceinline__ float exactValue(u2 w, float scale) {
    int32_t l = (int32_t)w.lo, h = (int32_t)w.hi;
    if ((h == 0 && l >= 0) || (h == -1 && l < 0)) return (float)l / scale;
    return ((float)h * 4294967296.0f + (float)w.lo) / scale;
}
__device__ __forceinline__ u2 addExact(u2 w, float x) {
    u2 p = exactParts(x, EXACT_FINE);
    u2 r;
    r.lo = w.lo + p.lo;
    r.hi = w.hi + p.hi + (r.lo < p.lo ? 1u : 0u);
    return r;
}

// the eight neighbour steps (common.wgsl:52-90)
__device__ __forceinline__ uint32_t stepIndex(i2 d) {
    if (d.y == 0) return d.x > 0 ? 0u : 4u;
    if (d.x == 0) return d.y > 0 ? 2u : 6u;
    if (d.y > 0) return d.x > 0 ? 1u : 3u;
    return d.x > 0 ? 7u : 5u;
}
__device__ __forceinline__ int32_t stepX(uint32_t i) {
    uint32_t j = (i + 1u) & 7u;
    return j <= 2u ? 1 : ((j >= 4u && j <= 6u) ? -1 : 0);
}
__device__ __forceinline__ i2 stepAt(uint32_t i) { return I2(stepX(i), stepX((i + 6u) & 7u)); }
__device__ __forceinline__ i2 turnStep(i2 d, int32_t k) { return stepAt((stepIndex(d) + (uint32_t)(k + 8)) & 7u); }

// water's effTempIn at 1 atm, expandedAt, liquidDensity, liquidVolume
// (common.wgsl:981-990,1501-1530,1560-1564; fluid.wgsl:1048-1052)
__device__ __forceinline__ float effTemp(float stored) {
    return wclamp(stored, WATER_LOT, WATER_HIT) + fmaxf(stored - WATER_HIT - WATER_UPSPAN, 0.0f)
        - fmaxf(WATER_LOT - WATER_DOWNSPAN - stored, 0.0f) * WATER_DOWNSCALE;
}
__device__ __forceinline__ float expandedAt(float temp) {
    float d = temp - fmaxf(AMBIENT_T, WATER_LOT);
    return wclamp(1.0f - WATER_EXPANSION * d - WATER_EXPANSION2 * d * d, 0.8f, 1.05f);
}
__device__ __forceinline__ float liquidDensity(float temp) { return 100.0f / 100.0f * expandedAt(effTemp(temp)); }
__device__ __forceinline__ float liquidVolume(float temp) { return 1.0f / expandedAt(effTemp(temp)); }

// ---------------------------------------------------------------- addressing
// Memory by address space.  The Dev struct's pointers are generic, and so
// are pointers into LDS once stored in a struct; an access through a generic
// pointer compiles to a FLAT instruction, which the LDS serves far slower
// than its own instructions and whose waits tie the LDS and scalar counters
// to the vector memory one.  These name the space: lds*() for pointers into
// __shared__ arrays, glb*() for device memory.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) const uint8_t lds_cu8;
typedef __attribute__((address_space(1))) uint32_t glb_u32;
__device__ __forceinline__ void ldsAdd(uint32_t *p, uint32_t v) {
    __hip_atomic_fetch_add((lds_u32 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t ldsGet(const uint32_t *p) { return *(const lds_u32 *)p; }
__device__ __forceinline__ uint8_t ldsGetU8(const uint8_t *p) { return *(lds_cu8 *)p; }
__device__ __forceinline__ void glbAdd(uint32_t *p, uint32_t v) {
    __hip_atomic_fetch_add((glb_u32 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// a device-memory object through a generic pointer, as device memory
template <typename T> __device__ __forceinline__ __attribute__((address_space(1))) T &G(T &x) {
    return *(__attribute__((address_space(1))) T *)&x;
}
#define CNT(d, i, g) ((d).counters[LIQUID_CELL_WORDS * (uint64_t)(i) + (g)])
#define SUM(d, f, k) ((d).counters[LIQUID_CELL_WORDS * (uint64_t)(d).C + 2ull * (f) + (k)])
#define FLD(d, v, c) ((d).field[4ull * (uint64_t)(v) + (c)])
#define CGV(d, i, c) ((d).cg[4ull * (uint64_t)(i) + (c)])
#define MOTION(d, pb, k, c) ((pb)[4u + 4ull * (uint64_t)(k) + (c)])
#define STATE(d, pb, k, c) ((pb)[4u + 4ull * ((uint64_t)(d).cap + (k)) + (c)])
#define AFFINE(d, pb, k, c) ((pb)[4u + 4ull * (2ull * (d).cap + (k)) + (c)])
#define STATE_SLOT(d, n) ((d).B + (n))
#define ACTIVE_SLOT(d, n) (2u * (d).B + (n))
#define RETIRED_SLOT(d, n) (3u * (d).B + (n))
#define COUNT_SLOT(d, k) (4u * (d).B + (k))
#define TOTAL_SLOT(d, n) (4u * (d).B + FLUID_STATS + (n))

__device__ __forceinline__ uint32_t kindOf(const Dev &d, uint32_t t) { return d.kind[t & 0xffu]; }
__device__ __forceinline__ bool inWorld(const Dev &d, i2 c) { return c.x >= 0 && c.y >= 0 && c.x < d.W && c.y < d.H; }
// (a cell clamped into a w x h grid: an address that is always there, for
// loads made whatever the test, their values then chosen)
__device__ __forceinline__ i2 clampTo(i2 c, int32_t w, int32_t h) {
    return I2(c.x < 0 ? 0 : (c.x >= w ? w - 1 : c.x), c.y < 0 ? 0 : (c.y >= h ? h - 1 : c.y));
}
__device__ __forceinline__ uint32_t cellIndex(const Dev &d, i2 c) { return (uint32_t)c.y * (uint32_t)d.W + (uint32_t)c.x; }
// The solve's own arrays (the packed fine level, the FP64 check's vectors)
// are block-major: each 8 x 8 block's cells contiguous, row-major inside it,
// so a wavefront working a block reads whole cache lines.  c in the world.
__device__ __forceinline__ uint32_t packIndex(const Dev &d, i2 c) {
    return (((uint32_t)c.y >> 3) * d.bw + ((uint32_t)c.x >> 3)) * 64u + (((uint32_t)c.y & 7u) << 3) + ((uint32_t)c.x & 7u);
}
__device__ __forceinline__ i2 cellOf(const Dev &d, v2 p) {
    int32_t x = wi32(floorf(p.x)), y = wi32(floorf(p.y));
    x = x < 0 ? 0 : (x > d.W - 1 ? d.W - 1 : x);
    y = y < 0 ? 0 : (y > d.H - 1 ? d.H - 1 : y);
    return I2(x, y);
}
__device__ __forceinline__ i2 beside(i2 c, int k) {
    return k < 2 ? I2(c.x + (k & 1) * 2 - 1, c.y) : I2(c.x, c.y + (k & 1) * 2 - 1);
}
__device__ __forceinline__ uint32_t uFace(const Dev &d, int32_t x, int32_t y) { return (uint32_t)y * (uint32_t)(d.W + 1) + (uint32_t)x; }
__device__ __forceinline__ uint32_t vFace(const Dev &d, int32_t x, int32_t y) {
    return (uint32_t)(d.W + 1) * (uint32_t)d.H + (uint32_t)y * (uint32_t)d.W + (uint32_t)x;
}
__device__ __forceinline__ bool faceInWorld(const Dev &d, i2 at, bool isU) {
    int32_t ex = d.W + (isU ? 1 : 0), ey = d.H + (isU ? 0 : 1);
    return at.x >= 0 && at.y >= 0 && at.x < ex && at.y < ey;
}
__device__ __forceinline__ uint32_t faceIndex(const Dev &d, i2 at, bool isU) { return isU ? uFace(d, at.x, at.y) : vFace(d, at.x, at.y); }
__device__ __forceinline__ uint64_t faceSlot(const Dev &d, uint32_t f) { return 2ull * d.C + f; }
__device__ __forceinline__ uint64_t viscRhsSlot(const Dev &d, uint32_t f) { return 2ull * d.C + d.F + f / 4u; }
__device__ __forceinline__ uint64_t viscLastSlot(const Dev &d, uint32_t f) { return 2ull * d.C + d.F + (d.F + 3u) / 4u + f / 4u; }

__device__ __forceinline__ bool inBlocks(const Dev &d, i2 b) { return b.x >= 0 && b.y >= 0 && b.x < (int32_t)d.bw && b.y * (int32_t)d.bw < (int32_t)d.B; }
__device__ __forceinline__ uint32_t blockIndex(const Dev &d, i2 b) { return (uint32_t)b.y * d.bw + (uint32_t)b.x; }
__device__ __forceinline__ uint32_t activeCount(const Dev &d) { return d.blocks[COUNT_SLOT(d, COUNT_ACTIVE)]; }
__device__ __forceinline__ bool isActiveBlock(const Dev &d, i2 b) { return inBlocks(d, b) && (d.blocks[STATE_SLOT(d, blockIndex(d, b))] & ACTIVE) != 0u; }
__device__ __forceinline__ i2 activeBlock(const Dev &d, uint32_t n) {
    uint32_t b = d.blocks[ACTIVE_SLOT(d, n)];
    return I2((int32_t)(b % d.bw), (int32_t)(b / d.bw));
}
__device__ __forceinline__ i2 idiv8(i2 c) { return I2(c.x / WIG_BLOCK, c.y / WIG_BLOCK); }

// the gas plane (common.wgsl:412-465)
__device__ __forceinline__ uint64_t gasWord(const Dev &d, uint32_t i) { return 4ull * (d.C + (i >> 1)) + ((i & 1u) ? 2u : 0u); }
